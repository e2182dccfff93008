// C ABI glue: argument validation, handle lifetime, error text, and the
// library's allocation wrappers.
#include <cstdarg>
#include <mutex>
#include <cstring>

#include "common.hpp"

namespace cosma_amd {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    return COSMA_AMD_ERR_HIP;
}

// The stream argument is used as given; NULL is the device's null stream,
// as for a BLAS handle whose stream was never set.
static hipStream_t pick_stream(cosma_amd_handle_t, void *stream) {
    return static_cast<hipStream_t>(stream);
}

// Checks the usual BLAS leading-dimension rule ld >= max(1, rows).
static bool ld_ok(int64_t ld, int64_t rows) { return ld >= (rows > 1 ? rows : 1); }

// Failure injection (cosma_amd_test_fail_allocations): allocations to let
// through first, then allocations to fail.
static std::mutex g_fail_mu;
static int g_fail_skip = 0, g_fail_count = 0;

static bool fail_this_allocation() {
    std::lock_guard<std::mutex> lock(g_fail_mu);
    if (g_fail_count == 0) return false;
    if (g_fail_skip > 0) {
        --g_fail_skip;
        return false;
    }
    --g_fail_count;
    return true;
}

hipError_t dev_alloc(void **p, size_t bytes) {
    if (fail_this_allocation()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

hipError_t pinned_alloc(void **p, size_t bytes) {
    if (fail_this_allocation()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

}  // namespace cosma_amd

using namespace cosma_amd;

extern "C" {

const char *cosma_amd_last_error(void) { return g_err; }

const char *cosma_amd_version(void) { return "cosma_amd 0.1 gfx950"; }

int cosma_amd_handle_create(cosma_amd_handle_t *handle, int device, int streams,
                            int64_t tile_m, int64_t tile_n, int64_t tile_k) {
    clear_error();
    CA_CHECK(handle != nullptr, "handle_create: null handle pointer");
    *handle = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("handle_create: no HIP device visible");
        return COSMA_AMD_ERR_NO_DEVICE;
    }
    CA_CHECK(device >= 0 && device < count, "handle_create: device %d out of range [0,%d)", device, count);
    hipDeviceProp_t prop;
    CA_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("handle_create: device %d is %s, kernels are built for gfx950", device, prop.gcnArchName);
        return COSMA_AMD_ERR_NO_DEVICE;
    }
    CA_HIP(hipSetDevice(device));
    // the no-allocation scratch of this device, resolved with the device
    // current; if that fails, only the no-allocation entry point refuses to run
    void *scratch_addr = nullptr;
    if (ozaki_scratch_address(device, &scratch_addr) != COSMA_AMD_OK) {
        scratch_addr = nullptr;
        (void)hipGetLastError();
        clear_error();
    }
    auto *h = new cosma_amd_handle_s();
    h->ozaki_scratch = scratch_addr;
    h->device = device;
    h->n_streams = streams > 0 ? streams : 2;
    h->tile_m = tile_m > 0 ? tile_m : 5000;
    h->tile_n = tile_n > 0 ? tile_n : 5000;
    h->tile_k = tile_k > 0 ? tile_k : 5000;
    h->n_cu = prop.multiProcessorCount;
    for (int i = 0; i < h->n_streams + 2; ++i) {
        hipStream_t s;
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (auto x : h->streams) hipStreamDestroy(x);
            if (h->h2d) hipStreamDestroy(h->h2d);
            delete h;
            return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
        }
        if (i < h->n_streams) h->streams.push_back(s);
        else if (!h->h2d) h->h2d = s;
        else h->d2h = s;
    }
    *handle = h;
    return COSMA_AMD_OK;
}

int cosma_amd_handle_get_config(cosma_amd_handle_t h, int *device, int *streams, int64_t *tile_m, int64_t *tile_n,
                                int64_t *tile_k) {
    clear_error();
    CA_CHECK(h != nullptr, "handle_get_config: null handle");
    if (device) *device = h->device;
    if (streams) *streams = h->n_streams;
    if (tile_m) *tile_m = h->tile_m;
    if (tile_n) *tile_n = h->tile_n;
    if (tile_k) *tile_k = h->tile_k;
    return COSMA_AMD_OK;
}

int cosma_amd_handle_destroy(cosma_amd_handle_t h) {
    clear_error();
    if (!h) return COSMA_AMD_OK;
    hipSetDevice(h->device);
    for (auto s : h->streams) {
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
    }
    for (auto s : {h->h2d, h->d2h}) {
        if (!s) continue;
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
    }
    release_stream_state(h);
    delete h;
    return COSMA_AMD_OK;
}

int cosma_amd_dgemm(cosma_amd_handle_t h, char transa, char transb,
                    int64_t m, int64_t n, int64_t k, double alpha,
                    const double *a, int64_t lda, const double *b, int64_t ldb,
                    double beta, double *c, int64_t ldc, void *stream) {
    clear_error();
    CA_CHECK(h != nullptr, "dgemm: null handle");
    const char ta = norm_trans(transa), tb = norm_trans(transb);
    CA_CHECK(ta != 0, "dgemm: invalid transa '%c'", transa);
    CA_CHECK(tb != 0, "dgemm: invalid transb '%c'", transb);
    CA_CHECK(m >= 0 && n >= 0 && k >= 0, "dgemm: negative extent m=%lld n=%lld k=%lld",
             (long long)m, (long long)n, (long long)k);
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    CA_CHECK(ld_ok(ldc, m), "dgemm: ldc=%lld < m=%lld", (long long)ldc, (long long)m);
    CA_CHECK(c != nullptr, "dgemm: null C");
    if (k > 0 && alpha != 0.0) {
        CA_CHECK(ld_ok(lda, ta == 'N' ? m : k), "dgemm: lda=%lld too small", (long long)lda);
        CA_CHECK(ld_ok(ldb, tb == 'N' ? k : n), "dgemm: ldb=%lld too small", (long long)ldb);
        CA_CHECK(a != nullptr && b != nullptr, "dgemm: null A or B");
    }
    CA_HIP(hipSetDevice(h->device));
    return launch_dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, pick_stream(h, stream));
}

int cosma_amd_dgemm_host(cosma_amd_handle_t h, char transa, char transb,
                         int64_t m, int64_t n, int64_t k, double alpha,
                         const double *a, int64_t lda, const double *b, int64_t ldb,
                         double beta, double *c, int64_t ldc, int pin_host_buffers, int copy_c_back) {
    (void)pin_host_buffers;   // pageable memory is always staged, never registered in place
    clear_error();
    CA_CHECK(h != nullptr, "dgemm_host: null handle");
    const char ta = norm_trans(transa), tb = norm_trans(transb);
    CA_CHECK(ta != 0, "dgemm_host: invalid transa '%c'", transa);
    CA_CHECK(tb != 0, "dgemm_host: invalid transb '%c'", transb);
    CA_CHECK(m >= 0 && n >= 0 && k >= 0, "dgemm_host: negative extent m=%lld n=%lld k=%lld",
             (long long)m, (long long)n, (long long)k);
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    CA_CHECK(ld_ok(ldc, m), "dgemm_host: ldc=%lld < m=%lld", (long long)ldc, (long long)m);
    CA_CHECK(c != nullptr, "dgemm_host: null C");
    if (k > 0 && alpha != 0.0) {
        CA_CHECK(ld_ok(lda, ta == 'N' ? m : k), "dgemm_host: lda=%lld too small", (long long)lda);
        CA_CHECK(ld_ok(ldb, tb == 'N' ? k : n), "dgemm_host: ldb=%lld too small", (long long)ldb);
        CA_CHECK(a != nullptr && b != nullptr, "dgemm_host: null A or B");
    }
    CA_HIP(hipSetDevice(h->device));
    return stream_dgemm(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, copy_c_back != 0);
}

static int dgemm_ozaki_entry(cosma_amd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                             double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                             double beta, double *c, int64_t ldc, int method, double tolerance, int count,
                             cosma_amd_ozaki_info *info, void *stream, bool noalloc);

int cosma_amd_dgemm_ozaki(cosma_amd_handle_t h, char transa, char transb,
                          int64_t m, int64_t n, int64_t k, double alpha,
                          const double *a, int64_t lda, const double *b, int64_t ldb,
                          double beta, double *c, int64_t ldc, int method, double tolerance,
                          int count, cosma_amd_ozaki_info *info, void *stream) {
    return dgemm_ozaki_entry(h, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, method, tolerance,
                             count, info, stream, false);
}

int cosma_amd_dgemm_ozaki_noalloc(cosma_amd_handle_t h, char transa, char transb,
                                  int64_t m, int64_t n, int64_t k, double alpha,
                                  const double *a, int64_t lda, const double *b, int64_t ldb,
                                  double beta, double *c, int64_t ldc, int method, double tolerance,
                                  int count, cosma_amd_ozaki_info *info, void *stream) {
    return dgemm_ozaki_entry(h, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, method, tolerance,
                             count, info, stream, true);
}

static int dgemm_ozaki_entry(cosma_amd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                             double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                             double beta, double *c, int64_t ldc, int method, double tolerance, int count,
                             cosma_amd_ozaki_info *info, void *stream, bool noalloc) {
    clear_error();
    CA_CHECK(h != nullptr, "dgemm_ozaki: null handle");
    const char ta = norm_trans(transa), tb = norm_trans(transb);
    CA_CHECK(ta != 0, "dgemm_ozaki: invalid transa '%c'", transa);
    CA_CHECK(tb != 0, "dgemm_ozaki: invalid transb '%c'", transb);
    const int base_method = method & ~COSMA_AMD_OZAKI_STRICT;
    CA_CHECK(base_method == COSMA_AMD_OZAKI_MODULI || base_method == COSMA_AMD_OZAKI_SLICES ||
                 base_method == COSMA_AMD_OZAKI_MODULI_BF16,
             "dgemm_ozaki: unknown method %d", method);
    CA_CHECK(m >= 0 && n >= 0 && k >= 0, "dgemm_ozaki: negative extent");
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    CA_CHECK(ld_ok(ldc, m), "dgemm_ozaki: ldc=%lld < m=%lld", (long long)ldc, (long long)m);
    CA_CHECK(c != nullptr, "dgemm_ozaki: null C");
    if (k > 0 && alpha != 0.0) {
        CA_CHECK(ld_ok(lda, ta == 'N' ? m : k), "dgemm_ozaki: lda=%lld too small", (long long)lda);
        CA_CHECK(ld_ok(ldb, tb == 'N' ? k : n), "dgemm_ozaki: ldb=%lld too small", (long long)ldb);
        CA_CHECK(a != nullptr && b != nullptr, "dgemm_ozaki: null A or B");
        // The INT8 / BF16 product kernel addresses a 256-row tile of an operand
        // plane with 32-bit offsets: 256 rows of round_up(k, 256) * (2 for
        // BF16) bytes must stay below 2^31.  Checked before any work is queued.
        const int64_t row_bytes = (k + 255) / 256 * 256 * (base_method == COSMA_AMD_OZAKI_MODULI_BF16 ? 2 : 1);
        if (row_bytes * 256 >= ((int64_t)1 << 31)) {
            set_error("dgemm_ozaki: k = %lld gives operand rows of %lld bytes; the product kernel supports rows "
                      "below %lld bytes", (long long)k, (long long)row_bytes, (long long)(((int64_t)1 << 31) / 256));
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
    }
    CA_HIP(hipSetDevice(h->device));
    if (noalloc)
        return ozaki_dgemm_noalloc(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, method, tolerance,
                                   count, info, pick_stream(h, stream));
    return ozaki_dgemm(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, method, tolerance, count,
                       info, pick_stream(h, stream));
}

int cosma_amd_get_device_c(cosma_amd_handle_t h, double **c, int64_t *ldc) {
    clear_error();
    CA_CHECK(h != nullptr && c != nullptr && ldc != nullptr, "get_device_c: null argument");
    *c = h->dev_c;
    *ldc = h->dev_c_ld;
    return COSMA_AMD_OK;
}

int cosma_amd_handle_set_memory_limit(cosma_amd_handle_t h, int64_t bytes) {
    clear_error();
    CA_CHECK(h != nullptr, "set_memory_limit: null handle");
    CA_CHECK(bytes >= 0, "set_memory_limit: negative limit");
    h->mem_limit = bytes;
    return COSMA_AMD_OK;
}

int cosma_amd_handle_last_mode(cosma_amd_handle_t h) { return h ? h->last_mode : -1; }

int cosma_amd_dgemm_set_tuning(int tile_config, int group_m) {
    clear_error();
    CA_CHECK(tile_config >= 0 && tile_config <= 4, "dgemm_set_tuning: tile_config %d not in 0..4", tile_config);
    CA_CHECK(group_m >= 0, "dgemm_set_tuning: negative group_m");
    set_dgemm_tuning(tile_config, group_m);
    return COSMA_AMD_OK;
}

int cosma_amd_block_transform(cosma_amd_handle_t h, char op, int64_t rows, int64_t cols,
                              double alpha, const double *src, int64_t ld_src, double beta,
                              double *dst, int64_t ld_dst, void *stream) {
    clear_error();
    CA_CHECK(h != nullptr, "block_transform: null handle");
    const char o = norm_trans(op);
    CA_CHECK(o != 0, "block_transform: invalid op '%c'", op);
    CA_CHECK(rows >= 0 && cols >= 0, "block_transform: negative extent");
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    CA_CHECK(src != nullptr && dst != nullptr, "block_transform: null pointer");
    CA_CHECK(ld_ok(ld_dst, rows), "block_transform: ld_dst=%lld < rows=%lld", (long long)ld_dst, (long long)rows);
    CA_CHECK(ld_ok(ld_src, o == 'N' ? rows : cols), "block_transform: ld_src=%lld too small", (long long)ld_src);
    CA_HIP(hipSetDevice(h->device));
    return launch_block_transform(o, rows, cols, alpha, src, ld_src, beta, dst, ld_dst, pick_stream(h, stream));
}

int cosma_amd_dgemm_layout(cosma_amd_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                           double alpha, const cosma_amd_block *a_blocks, int64_t n_a_blocks,
                           const cosma_amd_block *b_blocks, int64_t n_b_blocks, double beta,
                           const cosma_amd_block *c_blocks, int64_t n_c_blocks) {
    clear_error();
    CA_CHECK(h != nullptr, "dgemm_layout: null handle");
    const char ta = norm_trans(transa), tb = norm_trans(transb);
    CA_CHECK(ta != 0, "dgemm_layout: invalid transa '%c'", transa);
    CA_CHECK(tb != 0, "dgemm_layout: invalid transb '%c'", transb);
    CA_CHECK(m >= 0 && n >= 0 && k >= 0, "dgemm_layout: negative dimension");
    CA_HIP(hipSetDevice(h->device));
    return layout_dgemm(h, ta, tb, m, n, k, alpha, a_blocks, n_a_blocks, b_blocks, n_b_blocks, beta, c_blocks,
                        n_c_blocks);
}

int cosma_amd_grouped_create(cosma_amd_handle_t h, const cosma_amd_gemm_desc *problems, int64_t count,
                             int tile_config, cosma_amd_grouped_t *plan) {
    clear_error();
    CA_CHECK(h != nullptr, "grouped_create: null handle");
    CA_CHECK(plan != nullptr, "grouped_create: null plan pointer");
    CA_HIP(hipSetDevice(h->device));
    return grouped_create(h, problems, count, tile_config, plan);
}

int cosma_amd_grouped_run(cosma_amd_grouped_t plan, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_run: null plan");
    return grouped_run(plan, static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_graph_run(cosma_amd_grouped_t plan, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_graph_run: null plan");
    return grouped_graph_run(plan, static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_tune(cosma_amd_grouped_t plan, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_tune: null plan");
    return grouped_tune(plan, static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_set_config(cosma_amd_grouped_t plan, int tile_config, int grid) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_set_config: null plan");
    return grouped_set_config(plan, tile_config, grid);
}

int cosma_amd_grouped_info(cosma_amd_grouped_t plan, int *tile_config, int *grid, int64_t *tiles) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_info: null plan");
    grouped_info(plan, tile_config, grid, tiles);
    return COSMA_AMD_OK;
}

int cosma_amd_grouped_destroy(cosma_amd_grouped_t plan) {
    clear_error();
    grouped_destroy(plan);
    return COSMA_AMD_OK;
}

int cosma_amd_grouped_ozaki_create(cosma_amd_handle_t h, const cosma_amd_gemm_desc *problems, int64_t count,
                                   int moduli, double tolerance, cosma_amd_grouped_ozaki_t *plan) {
    return cosma_amd_grouped_ozaki_create_ex(h, problems, count, moduli, tolerance, COSMA_AMD_OZAKI_SPLIT_MIN_DIM,
                                             plan);
}

int cosma_amd_grouped_ozaki_create_ex(cosma_amd_handle_t h, const cosma_amd_gemm_desc *problems, int64_t count,
                                      int moduli, double tolerance, int64_t split_min_dim,
                                      cosma_amd_grouped_ozaki_t *plan) {
    clear_error();
    CA_CHECK(h != nullptr, "grouped_ozaki_create: null handle");
    CA_CHECK(plan != nullptr, "grouped_ozaki_create: null plan pointer");
    CA_HIP(hipSetDevice(h->device));
    return grouped_ozaki_create(h, problems, count, moduli, tolerance, split_min_dim, plan);
}

int cosma_amd_grouped_ozaki_run(cosma_amd_grouped_ozaki_t plan, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_run: null plan");
    return grouped_ozaki_run(plan, static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_ozaki_graph_run(cosma_amd_grouped_ozaki_t plan, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_graph_run: null plan");
    return grouped_ozaki_graph_run(plan, static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_ozaki_tune(cosma_amd_grouped_ozaki_t plan, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_tune: null plan");
    return grouped_ozaki_tune(plan, static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_ozaki_set_grid(cosma_amd_grouped_ozaki_t plan, int grid) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_set_grid: null plan");
    return grouped_ozaki_set_grid(plan, grid);
}

int cosma_amd_grouped_ozaki_info(cosma_amd_grouped_ozaki_t plan, int *moduli, double *error_estimate,
                                 int64_t *tiles, int *grid) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_info: null plan");
    grouped_ozaki_info(plan, moduli, error_estimate, tiles, grid);
    return COSMA_AMD_OK;
}

int cosma_amd_grouped_ozaki_routing(cosma_amd_grouped_ozaki_t plan, int64_t *split_problems,
                                    int64_t *fp64_problems) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_routing: null plan");
    grouped_ozaki_routing(plan, split_problems, fp64_problems);
    return COSMA_AMD_OK;
}

int cosma_amd_grouped_ozaki_check(cosma_amd_grouped_ozaki_t plan, double tolerance, double *worst_estimate,
                                  int64_t *worst_problem, int *moduli_needed, void *stream) {
    clear_error();
    CA_CHECK(plan != nullptr, "grouped_ozaki_check: null plan");
    CA_CHECK(worst_estimate && worst_problem && moduli_needed, "grouped_ozaki_check: null output");
    return grouped_ozaki_check(plan, tolerance, worst_estimate, worst_problem, moduli_needed,
                               static_cast<hipStream_t>(stream));
}

int cosma_amd_grouped_ozaki_destroy(cosma_amd_grouped_ozaki_t plan) {
    clear_error();
    grouped_ozaki_destroy(plan);
    return COSMA_AMD_OK;
}

int cosma_amd_host_alloc(size_t bytes, void **ptr) {
    clear_error();
    CA_CHECK(ptr != nullptr, "host_alloc: null output pointer");
    *ptr = nullptr;
    if (bytes == 0) return COSMA_AMD_OK;
    hipError_t e = pinned_alloc(ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        set_error("host_alloc: hipHostMalloc of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
        return COSMA_AMD_ERR_ALLOC;
    }
    return COSMA_AMD_OK;
}

int cosma_amd_host_free(void *ptr) {
    clear_error();
    if (ptr) CA_HIP(hipHostFree(ptr));
    return COSMA_AMD_OK;
}

int cosma_amd_fill_uniform(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                           uint64_t offset, void *stream) {
    clear_error();
    CA_CHECK(rows >= 0 && cols >= 0, "fill_uniform: negative extent");
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    CA_CHECK(x != nullptr, "fill_uniform: null pointer");
    CA_CHECK(ld_ok(ld, rows), "fill_uniform: ld=%lld < rows=%lld", (long long)ld, (long long)rows);
    return launch_fill_uniform(x, rows, cols, ld, seed, offset, static_cast<hipStream_t>(stream));
}

int cosma_amd_fill_lognormal(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, double phi,
                             void *stream) {
    clear_error();
    CA_CHECK(rows >= 0 && cols >= 0, "fill_lognormal: negative extent");
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    CA_CHECK(x != nullptr, "fill_lognormal: null pointer");
    CA_CHECK(ld_ok(ld, rows), "fill_lognormal: ld=%lld < rows=%lld", (long long)ld, (long long)rows);
    return launch_fill_lognormal(x, rows, cols, ld, seed, phi, static_cast<hipStream_t>(stream));
}

int cosma_amd_test_fail_allocations(int skip, int count) {
    clear_error();
    std::lock_guard<std::mutex> lock(g_fail_mu);
    g_fail_skip = skip > 0 ? skip : 0;
    g_fail_count = count > 0 ? count : 0;
    return COSMA_AMD_OK;
}

}  // extern "C"
