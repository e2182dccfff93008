/*
 * cosma_amd.h — C ABI of the MI355X (gfx950) local-multiply library.
 *
 * This is the thin layer between the C++ host code (cosma::local_multiply,
 * the pxgemm front-end) and the hand-written HIP kernels.  Plain C types only:
 * pointers, 64-bit extents, chars for transposes, int status codes.  Every
 * matrix is column-major.  A status of COSMA_AMD_OK (0) means success; any
 * other value is an error whose text is returned by cosma_amd_last_error().
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the COSMA v2.8.4 source tree; Tiled-MM and COSTA are external
 * libraries whose interfaces are known from COSMA's call sites).
 */
#ifndef COSMA_AMD_H
#define COSMA_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    COSMA_AMD_OK = 0,
    COSMA_AMD_ERR_INVALID_ARG = 1,  /* bad transpose char, negative extent, ld too small */
    COSMA_AMD_ERR_HIP = 2,          /* a HIP runtime call failed */
    COSMA_AMD_ERR_ALLOC = 3,        /* device or pinned allocation failed */
    COSMA_AMD_ERR_UNSUPPORTED = 4,  /* request outside what the kernels implement */
    COSMA_AMD_ERR_NO_DEVICE = 5     /* no gfx950 device visible */
};

/* Text of the last error raised on the calling thread ("" if none). */
const char *cosma_amd_last_error(void);

/* Library version string, e.g. "cosma_amd 0.1 gfx950". */
const char *cosma_amd_version(void);

/* ---------------------------------------------------------------- handle */
/*
 * Opaque handle owning the HIP streams, the host-streaming staging ring and
 * kernel workspace on one device.
 * Replaces Tiled-MM's gpu::mm_handle<T> and gpu::make_context<T>(streams,
 * tile_m, tile_n, tile_k) created by cosma_context
 * (src/cosma/context.cpp:24 and :45; declared via context.hpp:59).
 * streams <= 0 and tile_* <= 0 take the defaults of the reference
 * (COSMA_GPU_STREAMS=2, COSMA_GPU_MAX_TILE_{M,N,K}=5000,
 * src/cosma/environment_variables.hpp:57-62).
 */
typedef struct cosma_amd_handle_s *cosma_amd_handle_t;

int cosma_amd_handle_create(cosma_amd_handle_t *handle, int device, int streams,
                            int64_t tile_m, int64_t tile_n, int64_t tile_k);
int cosma_amd_handle_destroy(cosma_amd_handle_t handle);
/*
 * The device, compute-stream count and streaming tile sizes of a handle
 * (after the defaults were applied).  Diagnostic; Tiled-MM's handle keeps
 * the same settings, but COSMA's call sites never read them back.
 */
int cosma_amd_handle_get_config(cosma_amd_handle_t handle, int *device, int *streams, int64_t *tile_m,
                                int64_t *tile_n, int64_t *tile_k);

/* ----------------------------------------------------- device-resident GEMM */
/*
 * C <- alpha * op(A) * op(B) + beta * C on device-resident operands, FP64,
 * computed on the FP64 MFMA units (v_mfma_f64_16x16x4f64).
 * op(A) is m x k, op(B) is k x n, C is m x n.  transa/transb in {'N','T','C'}
 * (upper or lower case; 'C' equals 'T' for real data).
 * Semantics kept from the reference:
 *   m == 0 or n == 0  -> no-op;
 *   k == 0 or alpha == 0 -> C <- beta * C   (src/cosma/multiply.cpp:99-106,
 *                                            src/cosma/cosma_pxgemm.cpp:39-46);
 *   beta == 0 -> C is never read (NaN in C does not propagate;
 *                utils/pxgemm_utils.hpp:604-616).
 * stream is a hipStream_t (NULL = the null stream).  The call is
 * asynchronous with respect to the host.
 * Replaces the per-tile vendor GEMM that Tiled-MM's gpu::gemm issues
 * (call site src/cosma/local_multiply.cpp:245-260) and the host
 * cosma::gemm -> cblas_dgemm (src/cosma/blas.cpp:24-48).
 */
int cosma_amd_dgemm(cosma_amd_handle_t handle, char transa, char transb,
                    int64_t m, int64_t n, int64_t k, double alpha,
                    const double *a, int64_t lda, const double *b, int64_t ldb,
                    double beta, double *c, int64_t ldc, void *stream);

/* ------------------------------------------------ split-precision GEMM */
enum {
    COSMA_AMD_OZAKI_MODULI = 0,   /* integer residues modulo coprime moduli <= 256, CRT recombination */
    COSMA_AMD_OZAKI_SLICES = 1,   /* signed 7-bit slices, triangular slice products */
    COSMA_AMD_OZAKI_MODULI_BF16 = 2, /* the moduli method with BF16 residues: products on the BF16
                                        matrix cores (v_mfma_f32_16x16x32_bf16), exact FP32 sums
                                        reduced modulo each modulus every 896 k */
    COSMA_AMD_OZAKI_STRICT = 0x100 /* flag OR-ed into the method: the tolerance must hold for the
                                      rigorous error bound, not only for the estimate */
};

typedef struct {
    int method;             /* method used; -1 if A or B held Inf/NaN (FP64 MFMA DGEMM used instead) */
    int count;              /* moduli or slices used (after any raise by the check) */
    int bits;               /* bits kept per input entry relative to its row/column maximum (moduli: the
                               largest P for which the entrywise or the row/column-norm bound keeps |A'B'|
                               below half the moduli product; slices: 7 * count) */
    int64_t int8_products;  /* INT8 (BF16 for COSMA_AMD_OZAKI_MODULI_BF16) GEMMs of size m x n x k computed,
                               over all passes */
    double error_bound;     /* rigorous bound on ||X' - X||_F / ||X||_F, X = op(A) op(B) and X' the computed
                               product (before alpha and beta), from the rounding of every input entry and
                               the norm of X' (NaN for the FP64 fallback; +inf if X' is too small to bound) */
    double error_estimate;  /* the same ratio as expected with independent rounding errors */
    int tolerance_met;      /* 1: the criterion meets the tolerance (1.5 x the estimate, or the bound with
                               COSMA_AMD_OZAKI_STRICT); 0: it does not even at the count used (the largest
                               count, or a count given by the caller); -1: FP64 fallback */
    int passes;             /* product passes computed: 1, or more when the check raised the count */
    double product_ms;      /* device time of the INT8 / BF16 product launches, all passes */
} cosma_amd_ozaki_info;

/*
 * FP64-accurate C <- alpha * op(A) * op(B) + beta * C on device-resident
 * operands, computed on the INT8 matrix cores (v_mfma_i32_16x16x64_i8) with
 * exact INT32 accumulation.  Rows of op(A) and columns of op(B) are scaled by
 * powers of two (exponent of their maximum), the scaled entries are split into
 * INT8 operands (residues or slices, see the enum), the INT8 products are
 * computed, and FP64 recombines them and applies alpha and beta (beta == 0
 * never reads C).  FP64 arithmetic runs only in the scaling, the
 * recombination and the epilogue.
 * tolerance: requested relative accuracy (<= 0 means 1e-14) of the product
 * X = op(A) op(B) in the Frobenius norm, ||X' - X||_F <= tol ||X||_F.  The
 * count of moduli or slices is chosen from the data in two steps.  Before the
 * products, the row/column maxima and sums of squares of op(A) and op(B) give
 * an estimate of the error of each precision relative to the size X would
 * have with independent signs, and the smallest count for which twice the
 * estimate meets the tolerance is used.  After the products, the norm of the
 * computed X' (summed in the recombination) replaces that assumed size: when
 * cancellation makes X smaller and the error relative to it misses the
 * tolerance, the count is raised and the product recomputed (with beta != 0,
 * C is read and written only once, with the final count).  The check asks
 * 1.5 times the estimate to meet the tolerance, or with COSMA_AMD_OZAKI_STRICT
 * OR-ed into `method` the rigorous bound (from the rounding of every entry
 * and the row/column 1-norms).  count > 0 forces
 * the number of moduli (2..20) or slices (1..8) instead (no raise).  info
 * (may be NULL) receives the method, count, bits, the rigorous bound and the
 * estimate for the result, whether the tolerance was met, and the product
 * time.  Inputs whose error cannot meet the tolerance at the largest count
 * (strong cancellation: P is at most 58 bits below each row/column maximum)
 * are computed at that count and reported with tolerance_met = 0.
 * The BF16 variant (COSMA_AMD_OZAKI_MODULI_BF16) holds the same residues as
 * BF16 numbers (exact: integers in [-128, 127]) and multiplies them on the
 * BF16 matrix cores; it chooses the same count and gives bit-identical
 * results at twice the operand bytes.
 * Limits: moduli need k < 131072 (INT8 sums exact in INT32), the BF16 variant
 * needs round_up(k, 256) < 2^22 (32-bit offsets in the product kernel), and
 * slices need slices * k * 127^2 < 2^31; COSMA_AMD_ERR_UNSUPPORTED otherwise
 * (checked before any work is queued).
 * Uses workspaces owned by the handle; returns when C is final (one read-back
 * of the row statistics before the products, one of the product's norm per
 * pass).
 * No counterpart in the reference (its local multiply is a host or vendor
 * FP64 GEMM, src/cosma/local_multiply.cpp:299-381); same argument
 * conventions as cosma_amd_dgemm.
 */
int cosma_amd_dgemm_ozaki(cosma_amd_handle_t handle, char transa, char transb,
                          int64_t m, int64_t n, int64_t k, double alpha,
                          const double *a, int64_t lda, const double *b, int64_t ldb,
                          double beta, double *c, int64_t ldc, int method, double tolerance,
                          int count, cosma_amd_ozaki_info *info, void *stream);

/*
 * The same call with no allocation: no hipMalloc, hipFree, hipHostMalloc or
 * event creation happens in it, for callers that cannot allocate (tight
 * serving loops, harnesses that forbid device allocation).  No counterpart
 * in the reference.  The device workspace is a 64 MiB scratch array that
 * the library holds on every device it runs on until the process ends (its
 * address on each device is resolved when a handle is created there), and
 * the statistics are staged in a 1 MiB host buffer of the library.  Limit:
 * the statistics, residues / slices and product output of the largest count
 * the call may use (the forced count, else 20 moduli / 8 slices) must fit in
 * the 64 MiB; e.g. m = n = k = 1024 fits with any count.  Larger calls
 * return COSMA_AMD_ERR_UNSUPPORTED before any work is queued.  Results are
 * bit-identical to cosma_amd_dgemm_ozaki; info->product_ms is 0 (no timing
 * events).  Every call on a device shares that device's scratch (each call
 * clears what it reads), and the host staging is shared by the process:
 * no-allocation calls are serialised on the host, and calls that overlap on
 * one device through different streams race.
 */
int cosma_amd_dgemm_ozaki_noalloc(cosma_amd_handle_t handle, char transa, char transb,
                                  int64_t m, int64_t n, int64_t k, double alpha,
                                  const double *a, int64_t lda, const double *b, int64_t ldb,
                                  double beta, double *c, int64_t ldc, int method, double tolerance,
                                  int count, cosma_amd_ozaki_info *info, void *stream);

/* ------------------------------------------------- host-streaming GEMM */
/*
 * Same contract as cosma_amd_dgemm, but each of A, B, C may be device memory,
 * pinned host memory (hipHostMalloc) or pageable host memory; the call
 * returns when C is final (on the host if copy_c_back != 0).
 * C is cut into tile_m x tile_n tiles and k into tile_k panels (the handle's
 * tile sizes, i.e. COSMA_GPU_MAX_TILE_{M,N,K}); uploads, per-tile MFMA DGEMMs
 * on the handle's round-robin streams (COSMA_GPU_STREAMS) and downloads of
 * finished C tiles overlap.  When A, B and C fit in the device-memory budget
 * (cosma_amd_handle_set_memory_limit, default 80% of free HBM) every block
 * crosses the host link once; otherwise blocks are streamed through a
 * bounded ring of device slots.  Pageable memory is staged through pinned
 * buffers owned by the handle; pin_host_buffers is accepted for
 * compatibility and ignored (caller memory is never registered in place).
 * copy_c_back == 0 leaves C on the device (see cosma_amd_get_device_c).
 * Replaces Tiled-MM's gpu::gemm(mm_handle<T>&, char ta, char tb, int m, int n,
 * int k, T alpha, const T* a, int lda, const T* b, int ldb, T beta, T* c,
 * int ldc, bool pin_host_buffers, bool copy_c_back) as called from
 * src/cosma/local_multiply.cpp:245-260, with 64-bit extents.
 */
int cosma_amd_dgemm_host(cosma_amd_handle_t handle, char transa, char transb,
                         int64_t m, int64_t n, int64_t k, double alpha,
                         const double *a, int64_t lda, const double *b, int64_t ldb,
                         double beta, double *c, int64_t ldc, int pin_host_buffers,
                         int copy_c_back);

/*
 * Device copy of C left by the last cosma_amd_dgemm_host call with
 * copy_c_back == 0 (NULL otherwise), valid until the next streaming call.
 * Replaces mm_handle<T>::get_full_device_buffer_c() (call site
 * src/cosma/gpu/nccl_utils.cpp:199-201).
 */
int cosma_amd_get_device_c(cosma_amd_handle_t handle, double **c, int64_t *ldc);

/*
 * Device-memory budget of the host streamer in bytes (0 = 80% of free HBM).
 * Plays the role of the tile-count bound of Tiled-MM's device buffers
 * (streams * (tm*tk + tk*tn + tm*tn) elements); no direct counterpart.
 */
int cosma_amd_handle_set_memory_limit(cosma_amd_handle_t handle, int64_t bytes);

/*
 * Schedule used by the last cosma_amd_dgemm_host call: 0 all operands on the
 * device, 1 resident panels, 2 bounded ring; -1 for a null handle.
 * Diagnostic; no counterpart in the reference.
 */
int cosma_amd_handle_last_mode(cosma_amd_handle_t handle);

/*
 * Selects the DGEMM tile shape for later cosma_amd_dgemm calls (process-wide):
 * tile_config 0 = automatic, 1 = 128x128, 2 = 256x128, 3 = 128x256 (m x n of
 * C per workgroup), 4 = the kernel with direct LDS loads (128x128, two
 * workgroups per CU; m and n multiples of 128, k of 16, 16-byte aligned
 * operands with even leading dimensions; other calls take the automatic
 * choice); group_m = tile-rows walked before moving along n
 * (0 = default for the shape).  No counterpart in the reference (Tiled-MM's
 * tile sizes, COSMA_GPU_MAX_TILE_*, describe host streaming tiles, not
 * kernel tiles); used by the tuner and the benchmarks.
 */
int cosma_amd_dgemm_set_tuning(int tile_config, int group_m);

/* ------------------------------------------------ local block transform */
/*
 * dst <- alpha * op(src) + beta * dst for one column-major block on the
 * device; op in {'N','T','C'}; op(src) is rows x cols.  With alpha == 1 and
 * beta == 0 this is a pure permutation and the result is bit-exact.
 * beta == 0 never reads dst.
 * Replaces COSTA's local block copy/transpose
 * costa::memory::copy_and_transpose<T>(src, n_rows, n_cols, src_stride,
 * dest, dest_stride, conj, workspace) (call site benchmarks/transpose.cpp:46-47)
 * and the per-block alpha/beta update of costa::transformer<T>::schedule
 * (call sites src/cosma/cosma_pxgemm.cpp:338-364, src/cosma/multiply.cpp:186-204).
 * Here rows/cols describe op(src), i.e. dst is rows x cols.
 */
int cosma_amd_block_transform(cosma_amd_handle_t handle, char op,
                              int64_t rows, int64_t cols, double alpha,
                              const double *src, int64_t ld_src, double beta,
                              double *dst, int64_t ld_dst, void *stream);

/* ------------------------------------------------------- grouped GEMM */
/* One problem of a grouped GEMM: C <- alpha * op(A) * op(B) + beta * C. */
typedef struct {
    char transa, transb;
    int64_t m, n, k;
    double alpha;
    const double *a;
    int64_t lda;
    const double *b;
    int64_t ldb;
    double beta;
    double *c;
    int64_t ldc;
} cosma_amd_gemm_desc;

typedef struct cosma_amd_grouped_s *cosma_amd_grouped_t;

/*
 * Plans a grouped GEMM over `count` independent device-resident problems
 * (same semantics per problem as cosma_amd_dgemm; problems with m or n == 0
 * are dropped).  The descriptors are copied to the device; the operand
 * pointers must stay valid while the plan is used.  tile_config: 0 = the
 * tuned choice for these shapes if one exists in the process, else 64x64;
 * 1 = 64x64, 2 = 128x64, 3 = 128x128 C tiles staged through LDS by four
 * waves; 4 = 64x64 and 5 = 32x32 C tiles computed by one wave each with
 * fragments loaded straight from global memory (no LDS).
 * cosma_amd_grouped_run: one persistent kernel launch on `stream` (no
 * allocation, memset or synchronisation: capturable).
 * cosma_amd_grouped_graph_run: captures the launch into a hipGraph on first
 * use and launches the graph on `stream`.
 * cosma_amd_grouped_tune: times every tile shape with several persistent
 * grid sizes on the plan's own problems (overwriting their C), keeps the
 * fastest and records it for later plans with the same shapes.
 * Each problem's result is bit-identical to cosma_amd_dgemm on it.
 * No counterpart in the reference (its local multiplies are one GEMM each,
 * src/cosma/local_multiply.cpp:299-381); on one rank the sequential steps of
 * src/cosma/multiply.cpp:460-588 are one client.
 */
int cosma_amd_grouped_create(cosma_amd_handle_t handle, const cosma_amd_gemm_desc *problems,
                             int64_t count, int tile_config, cosma_amd_grouped_t *plan);
int cosma_amd_grouped_run(cosma_amd_grouped_t plan, void *stream);
int cosma_amd_grouped_graph_run(cosma_amd_grouped_t plan, void *stream);
int cosma_amd_grouped_tune(cosma_amd_grouped_t plan, void *stream);
/* Forces a tile configuration (1..5) and grid (<= 0: default) for the plan. */
int cosma_amd_grouped_set_config(cosma_amd_grouped_t plan, int tile_config, int grid);
/* Current tile configuration, grid and number of tiles of the plan. */
int cosma_amd_grouped_info(cosma_amd_grouped_t plan, int *tile_config, int *grid, int64_t *tiles);
int cosma_amd_grouped_destroy(cosma_amd_grouped_t plan);

/* ------------------------------------------- grouped split-precision GEMM */
typedef struct cosma_amd_grouped_ozaki_s *cosma_amd_grouped_ozaki_t;

/*
 * Problems of a grouped split-precision plan whose smallest dimension
 * min(m, n, k) is below this are computed by the FP64 MFMA grouped kernel
 * instead (see cosma_amd_grouped_ozaki_create_ex).
 */
#define COSMA_AMD_OZAKI_SPLIT_MIN_DIM 2048

/*
 * Plans the split-precision (moduli method, as cosma_amd_dgemm_ozaki) form of
 * a grouped GEMM over `count` independent device-resident problems: per run,
 * four grouped launches (exponents, residues, the INT8 products of every
 * problem and modulus in one persistent launch over 64 x 64 tiles,
 * recombination with the alpha / beta epilogue).  One moduli count serves the
 * plan: `moduli` (2..20), or 0 to choose the smallest count whose doubled
 * error estimate meets `tolerance` (<= 0: 1e-14) for every problem, estimated
 * on the operands present at creation.  Problems whose operands hold Inf or
 * NaN are computed by FP64 dot products.  The descriptors and a workspace
 * (residues, exponents) are allocated at creation; runs allocate nothing and
 * do not synchronise, so they can be captured.
 * cosma_amd_grouped_ozaki_graph_run: captures a run into a hipGraph on first
 * use and launches the graph on `stream`.
 * cosma_amd_grouped_ozaki_tune: times the run with persistent product grids
 * of 1, 2, 4, 8 and 16 workgroups per CU and one per tile (overwriting the
 * problems' C) and keeps the fastest.
 * cosma_amd_grouped_ozaki_info: moduli count, doubled error estimate of the
 * worst problem at creation, number of INT8 tiles, product grid.
 * No counterpart in the reference (its local multiplies are plain GEMMs,
 * src/cosma/local_multiply.cpp:299-381).
 */
int cosma_amd_grouped_ozaki_create(cosma_amd_handle_t handle, const cosma_amd_gemm_desc *problems,
                                   int64_t count, int moduli, double tolerance,
                                   cosma_amd_grouped_ozaki_t *plan);
/*
 * The same with an explicit split threshold: problems with min(m, n, k) <
 * split_min_dim are computed in the same run (and graph) by the FP64 MFMA
 * grouped kernel of cosma_amd_grouped_create, which below that size is
 * faster than splitting and FP64-accurate (the FP64 GEMM's error bound); the
 * rest by the split method.  cosma_amd_grouped_ozaki_create uses
 * COSMA_AMD_OZAKI_SPLIT_MIN_DIM; 0 splits every problem.  The moduli count,
 * its estimate and cosma_amd_grouped_ozaki_check cover the split problems
 * (a count of 0 is reported when none is split and none was given).
 */
int cosma_amd_grouped_ozaki_create_ex(cosma_amd_handle_t handle, const cosma_amd_gemm_desc *problems,
                                      int64_t count, int moduli, double tolerance, int64_t split_min_dim,
                                      cosma_amd_grouped_ozaki_t *plan);
/* Number of problems of the plan computed by the split method and in FP64. */
int cosma_amd_grouped_ozaki_routing(cosma_amd_grouped_ozaki_t plan, int64_t *split_problems,
                                    int64_t *fp64_problems);
int cosma_amd_grouped_ozaki_run(cosma_amd_grouped_ozaki_t plan, void *stream);
int cosma_amd_grouped_ozaki_graph_run(cosma_amd_grouped_ozaki_t plan, void *stream);
int cosma_amd_grouped_ozaki_tune(cosma_amd_grouped_ozaki_t plan, void *stream);
int cosma_amd_grouped_ozaki_set_grid(cosma_amd_grouped_ozaki_t plan, int grid);
int cosma_amd_grouped_ozaki_info(cosma_amd_grouped_ozaki_t plan, int *moduli, double *error_estimate,
                                 int64_t *tiles, int *grid);
/*
 * Checks the last run of the plan against `tolerance` (<= 0: 1e-14): for
 * every problem the expected relative error of its computed product X'
 * (before alpha and beta), with the norm of X' summed in that run's
 * recombination and the error factors of the operands now present.  Returns
 * the largest of these estimates, the caller's index of that problem (-1 if
 * no problem uses the moduli), and the smallest moduli count for which
 * 1.5 x every estimate meets the tolerance (MAX + 1 = 21 if none does).  The
 * plan's count is fixed (its graph is); when moduli_needed exceeds it, the
 * operands have changed (or cancel) more than the count chosen at creation
 * allows, and a plan created with moduli_needed computes them to the
 * tolerance.  Synchronises `stream`; call it after a run on the same stream.
 */
int cosma_amd_grouped_ozaki_check(cosma_amd_grouped_ozaki_t plan, double tolerance, double *worst_estimate,
                                  int64_t *worst_problem, int *moduli_needed, void *stream);
int cosma_amd_grouped_ozaki_destroy(cosma_amd_grouped_ozaki_t plan);

/* ------------------------------------------------------ failure injection */
/*
 * Lets the next `skip` device (hipMalloc) or page-locked host (hipHostMalloc)
 * allocations of the library through and makes the `count` after them fail
 * as if memory were exhausted (count 0 disarms), for tests of the error
 * paths: every entry point then returns COSMA_AMD_ERR_ALLOC with a message
 * and leaves its handle or plan usable.  Process-wide.  No counterpart in
 * the reference, whose memory pool prints a hint and re-throws
 * std::bad_alloc (src/cosma/memory_pool.cpp:65-89,130-150); for tests only.
 */
int cosma_amd_test_fail_allocations(int skip, int count);

/* ------------------------------------------------------ pinned host memory */
/*
 * Allocates `bytes` of page-locked host memory (hipHostMalloc) that the
 * host-streaming GEMM copies by DMA without staging; frees it.  The C++
 * front-end's CosmaMatrix stores its local matrix here.  Replaces the
 * reference's host memory pool plus in-place registration of its buffers
 * (memory_pool<T>::pin, src/cosma/memory_pool.cpp:153-165;
 * pinned_buffers<T>::add, src/cosma/pinned_buffers.cpp:6-47).
 */
int cosma_amd_host_alloc(size_t bytes, void **ptr);
int cosma_amd_host_free(void *ptr);

/* ----------------------------------------------- grid-layout multiply */
/*
 * One local block of a matrix given as a grid of blocks: the block covers
 * stored rows [row, row + rows) and columns [col, col + cols) of the matrix
 * and its element (i, j) is data[i + j * ld] (host or device memory, like the
 * operands of cosma_amd_dgemm_host).
 * Plays the role of COSTA's costa::block<T> built from the C API struct
 * `block` (src/cosma/cinterface.hpp:16-21, src/cosma/cinterface.cpp:20-31).
 */
typedef struct {
    void *data;
    int64_t ld;
    int64_t row, col;
    int64_t rows, cols;
} cosma_amd_block;

/*
 * C <- alpha * op(A) * op(B) + beta * C for matrices that are each given as a
 * set of blocks that tile the stored matrix exactly (A is m x k if transa is
 * 'N', else k x m; B likewise; C is m x n).  The blocks are gathered into
 * device copies (block uploads / the block-copy kernel), the product
 * X = op(A) * op(B) is formed once on the device with the FP64 MFMA DGEMM,
 * and every C block is updated as C_blk <- alpha * X_blk + beta * C_blk by
 * the block-transform kernel (the reference's order: multiply with (1, 0),
 * then the layout transform applies alpha and beta).  beta == 0 never reads
 * C; k == 0 or alpha == 0 gives beta * C.  Blocks that overlap, leave a gap,
 * or fall outside the matrix are rejected.  Returns when C is final.
 * Replaces the single-rank work of cosma::multiply_using_layout<T>
 * (src/cosma/multiply.cpp:77-213) behind dmultiply_using_layout
 * (src/cosma/cinterface.hpp:52-59).
 */
int cosma_amd_dgemm_layout(cosma_amd_handle_t handle, char transa, char transb,
                           int64_t m, int64_t n, int64_t k, double alpha,
                           const cosma_amd_block *a_blocks, int64_t n_a_blocks,
                           const cosma_amd_block *b_blocks, int64_t n_b_blocks,
                           double beta, const cosma_amd_block *c_blocks,
                           int64_t n_c_blocks);

/* ------------------------------------------------------- synthetic inputs */
/*
 * Fills a column-major rows x cols device matrix (leading dimension ld) with
 * uniform [-1, 1) values: element (r, c) = u(splitmix64(seed *
 * 0x9E3779B97F4A7C15 + offset + r + c * rows)), u(x) = ((x >> 11) * 2^-53) * 2 - 1,
 * exact in FP64 and reproducible on the host.
 * Replaces the test generator of src/cosma/random_generator.hpp:26-32 (seeded
 * from std::random_device, not replayable) and the miniapp fill
 * (miniapp/cosma_miniapp.cpp:21-25).
 */
int cosma_amd_fill_uniform(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                           uint64_t offset, void *stream);

/*
 * Fills a column-major device matrix with a_ij = (u - 0.5) * exp(phi * z),
 * u ~ U[0,1), z ~ N(0,1) (Box-Muller), from the same counter-based stream
 * (element i uses positions 3i..3i+2).  phi controls the spread of exponents
 * within rows and columns, the hard case for split-precision GEMM (phi = 0:
 * uniform; phi = 4: entries spread over ~2^+-17).  Synthetic-input utility
 * for the split-precision benchmark; no counterpart in the reference.
 */
int cosma_amd_fill_lognormal(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                             double phi, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* COSMA_AMD_H */
