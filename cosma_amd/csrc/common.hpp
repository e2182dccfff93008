// Shared definitions for the gfx950 kernels and the C ABI glue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cosma_amd.h"

namespace cosma_amd {

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Thread-local error text backing cosma_amd_last_error().
void set_error(const char *fmt, ...);
// Every device (hipMalloc) and page-locked host (hipHostMalloc) allocation of
// the library goes through these, so that tests can make the next ones fail
// (cosma_amd_test_fail_allocations).
hipError_t dev_alloc(void **p, size_t bytes);
hipError_t pinned_alloc(void **p, size_t bytes);
void clear_error();

// Records a HIP failure and returns COSMA_AMD_ERR_HIP.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define CA_HIP(call)                                                      \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess)                                             \
            return ::cosma_amd::hip_fail(e_, #call, __FILE__, __LINE__);  \
    } while (0)

#define CA_CHECK(cond, ...)                                               \
    do {                                                                  \
        if (!(cond)) {                                                    \
            ::cosma_amd::set_error(__VA_ARGS__);                          \
            return COSMA_AMD_ERR_INVALID_ARG;                             \
        }                                                                 \
    } while (0)

// Normalises a BLAS transpose character; returns 'N', 'T' or 0 if invalid.
// 'C' (conjugate transpose) equals 'T' for real data.
inline char norm_trans(char t) {
    switch (t) {
    case 'N': case 'n': return 'N';
    case 'T': case 't': case 'C': case 'c': return 'T';
    default: return 0;
    }
}

}  // namespace cosma_amd

namespace cosma_amd {
// A device allocation that grows on demand and is reused across calls.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};
// A pinned host staging slot with the event of its last transfer.
struct StagingSlot {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    // pending copy-out from this slot to pageable host memory (D2H slots)
    double *out = nullptr;
    int64_t out_ld = 0, rows = 0, cols = 0;
};
}  // namespace cosma_amd

// Handle: one device, its streams and lazily grown workspaces.
struct cosma_amd_handle_s {
    int device = 0;
    int n_streams = 2;
    int64_t tile_m = 5000, tile_n = 5000, tile_k = 5000;
    std::vector<hipStream_t> streams;   // compute streams (round-robin)
    hipStream_t h2d = nullptr;          // host -> device copies
    hipStream_t d2h = nullptr;          // device -> host copies
    int n_cu = 256;
    // host-streaming state
    int64_t mem_limit = 0;              // device bytes the streamer may use (0 = 80% of free)
    cosma_amd::DevBuf dA, dB, dC;       // resident operand copies / ring slots
    cosma_amd::DevBuf dX;               // product before the layout epilogue
    cosma_amd::DevBuf dC0;              // ring of incoming C tiles (deferred beta * C)
    cosma_amd::DevBuf ozaki_ws;         // split-precision workspace
    cosma_amd::DevBuf ozaki_aux;        // split-precision row statistics and exponents
    cosma_amd::DevBuf ozaki_host;       // page-locked host copy of the statistics (hipHostMalloc)
    hipEvent_t ozaki_ev[2] = {nullptr, nullptr};   // around the split-precision product launch
    void *ozaki_scratch = nullptr;      // device address of the no-allocation scratch array on this device
    std::vector<hipEvent_t> events;     // event pool
    size_t events_used = 0;
    std::vector<cosma_amd::StagingSlot> h2d_slots, d2h_slots;
    size_t h2d_next = 0, d2h_next = 0;
    double *dev_c = nullptr;            // result left on the device (copy_c_back == 0)
    int64_t dev_c_ld = 0;
    int last_mode = 0;                  // 0 = device, 1 = resident panels, 2 = bounded ring
};

namespace cosma_amd {
// One problem of a grouped GEMM as the device sees it.
struct GroupedProblem {
    int64_t m, n, k;
    double alpha, beta;
    const double *a;
    int64_t lda;
    const double *b;
    int64_t ldb;
    double *c;
    int64_t ldc;
    int64_t tile_start;   // index of the problem's first tile in the launch
    int flags;            // bit 0: A transposed, bit 1: B transposed, bit 2: 16-byte operand loads
    int vec_c;            // 16-byte C accesses
};
static_assert(sizeof(GroupedProblem) % 8 == 0, "the int64 tile starts follow the descriptor array");
int grouped_tile_shape(int cfg, int *bm, int *bn);
int grouped_problem_flags(char ta, char tb, int64_t m, int64_t n, const double *a, int64_t lda, const double *b,
                          int64_t ldb, const double *c, int64_t ldc, int *flags, int *vec_c);
// probs: nprob descriptors in device memory, sorted as planned, followed by
// their nprob int64 tile_start values (the kernels search that array).
int launch_grouped(int cfg, const GroupedProblem *probs, int nprob, int64_t total_tiles, int grid, hipStream_t s);
int grouped_create(cosma_amd_handle_t h, const cosma_amd_gemm_desc *d, int64_t count, int tile_config,
                   cosma_amd_grouped_t *out);
int grouped_ozaki_create(cosma_amd_handle_t h, const cosma_amd_gemm_desc *d, int64_t count, int moduli,
                         double tol, int64_t split_min, cosma_amd_grouped_ozaki_t *out);
void grouped_ozaki_routing(cosma_amd_grouped_ozaki_t p, int64_t *split, int64_t *fp64);
int grouped_ozaki_run(cosma_amd_grouped_ozaki_t p, hipStream_t s);
int grouped_ozaki_graph_run(cosma_amd_grouped_ozaki_t p, hipStream_t s);
int grouped_ozaki_tune(cosma_amd_grouped_ozaki_t p, hipStream_t s);
int grouped_ozaki_set_grid(cosma_amd_grouped_ozaki_t p, int grid);
void grouped_ozaki_info(cosma_amd_grouped_ozaki_t p, int *moduli, double *estimate, int64_t *tiles, int *grid);
int grouped_ozaki_check(cosma_amd_grouped_ozaki_t p, double tol, double *worst, int64_t *worst_problem,
                        int *moduli_needed, hipStream_t s);
int grouped_ozaki_destroy(cosma_amd_grouped_ozaki_t p);
int grouped_run(cosma_amd_grouped_t p, hipStream_t s);
int grouped_graph_run(cosma_amd_grouped_t p, hipStream_t s);
int grouped_tune(cosma_amd_grouped_t p, hipStream_t s);
int grouped_set_config(cosma_amd_grouped_t p, int tile_config, int grid);
void grouped_info(cosma_amd_grouped_t p, int *tile_config, int *grid, int64_t *tiles);
void grouped_destroy(cosma_amd_grouped_t p);
}  // namespace cosma_amd

namespace cosma_amd {
// Kernel launchers (stream already resolved, arguments already validated).
int launch_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double *a, int64_t lda, const double *b, int64_t ldb,
                 double beta, double *c, int64_t ldc, hipStream_t s);
int launch_dgemm_cfg(int cfg, int group_m, char ta, char tb, int64_t m, int64_t n, int64_t k,
                     double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                     double beta, double *c, int64_t ldc, hipStream_t s);
void set_dgemm_tuning(int cfg, int group_m);
int launch_scale(int64_t m, int64_t n, double beta, double *c, int64_t ldc,
                 hipStream_t s);
int launch_add_scaled(int64_t m, int64_t n, double beta, const double *x, int64_t ldx, double *c,
                      int64_t ldc, hipStream_t s);
int launch_fill_uniform(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                        uint64_t offset, hipStream_t s);
int stream_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                 double beta, double *c, int64_t ldc, bool copy_c_back);
void release_stream_state(cosma_amd_handle_t h);
int layout_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const cosma_amd_block *A, int64_t na, const cosma_amd_block *B, int64_t nb, double beta,
                 const cosma_amd_block *C, int64_t nc);
// Caller-provided memory for a split-precision call that allocates nothing:
// device scratch for the statistics and the residues / slices, and host
// staging for the statistics read-back.
struct OzakiFixedMemory {
    void *device;
    size_t device_bytes;
    double *host;
    size_t host_bytes;
};
int ozaki_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                int64_t ldc, int method, double tol, int count, cosma_amd_ozaki_info *info, hipStream_t s,
                const OzakiFixedMemory *fixed = nullptr);
// Device address of the split-precision scratch array on `device` (resolved
// once per device, with that device current); the handle keeps it.
int ozaki_scratch_address(int device, void **ptr);
// The no-allocation split-precision call: ozaki_dgemm on the device's
// scratch array and a process-wide host staging buffer (serialised).
int ozaki_dgemm_noalloc(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                        int64_t ldc, int method, double tol, int count, cosma_amd_ozaki_info *info,
                        hipStream_t s);
int moduli_bits(int n, int64_t k);
int launch_fill_lognormal(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, double phi,
                          hipStream_t s);
int launch_block_transform(char op, int64_t rows, int64_t cols, double alpha,
                           const double *src, int64_t lds, double beta,
                           double *dst, int64_t ldd, hipStream_t s);
}  // namespace cosma_amd
