// ScaLAPACK-layout GEMM on one rank: cosma::pxgemm<double>, scale_matrix,
// the descriptor helpers and the COSMA-prefixed C entry points.
//
// Reference: src/cosma/cosma_pxgemm.cpp:15-460,652-655 (pxgemm, scale_matrix,
// is_problem_too_small), src/cosma/scalapack.cpp:1-143 (descriptor helpers,
// numroc), src/cosma/prefixed_pxgemm.cpp (cosma_p?gemm entry points).
#include <algorithm>
#include <cctype>
#include <string>

#include <cosma/context.hpp>
#include <cosma/cosma_pxgemm.hpp>
#include <cosma/environment_variables.hpp>
#include <cosma/prefixed_pxgemm.h>
#include <cosma/pxgemm.h>

#include "front.hpp"

namespace cosma {
namespace scalapack {

grid_info get_grid_info(int ctxt) {
    grid_info g;
    using gridinfo_fn = void (*)(int, int *, int *, int *, int *);
    auto f = reinterpret_cast<gridinfo_fn>(detail::find_symbol("Cblacs_gridinfo"));
    if (f) f(ctxt, &g.nprow, &g.npcol, &g.myrow, &g.mycol);
    return g;
}

int get_grid_context(const int *desca, const int *descb, const int *descc) {
    if (desca[1] != descb[1] || descb[1] != descc[1])
        detail::fail("pxgemm: the descriptors of A, B and C belong to different BLACS contexts");
    return desca[1];
}

int get_grid_context(const int *desc) { return desc[1]; }

int leading_dimension(const int *desc) { return desc[8]; }

int numroc(int n, int nb, int proc_coord, int proc_src, int n_procs) {
    // whole blocks, dealt round-robin from proc_src; the last, partial block
    // goes to the process after the last whole one
    const int blocks = n / nb;
    const int offset = (n_procs + proc_coord - proc_src) % n_procs;
    int owned = (blocks / n_procs) * nb;
    const int extra = blocks % n_procs;
    if (offset < extra) owned += nb;
    if (offset == extra) owned += n % nb;
    return owned;
}

int min_leading_dimension(int n, int nb, int rank_grid_dim) { return (n / nb / rank_grid_dim) * nb; }

int max_leading_dimension(int n, int nb, int rank_grid_dim) {
    const int blocks = n / nb;
    return min_leading_dimension(n, nb, rank_grid_dim) + ((blocks % rank_grid_dim == 0) ? n % nb : nb);
}

int local_buffer_size(const int *desc) {
    const grid_info g = get_grid_info(desc[1]);
    return leading_dimension(desc) * numroc(desc[3], desc[5], g.mycol, desc[7], g.npcol);
}

}  // namespace scalapack

bool is_problem_too_small(int m, int n, int k) {
    static const int threshold = get_cosma_dim_threshold();
    return std::min(m, std::min(n, k)) < threshold;
}

namespace {

// Checks one descriptor for a submatrix of `rows` x `cols` starting at
// (i, j) (1-based) on a 1 x 1 grid; returns the offset of its first element.
long long submatrix_offset(const char *name, const int *desc, int i, int j, int rows, int cols) {
    const std::string who = std::string("pxgemm: ") + name;
    if (desc[0] != 1) detail::fail(who + ": descriptor type " + std::to_string(desc[0]) + " is not 1 (dense)");
    if (desc[4] < 1 || desc[5] < 1) detail::fail(who + ": block sizes must be positive");
    if (desc[6] != 0 || desc[7] != 0)
        detail::fail(who + ": source process (" + std::to_string(desc[6]) + ", " + std::to_string(desc[7]) +
                     ") does not exist on a 1 x 1 grid");
    if (i < 1 || j < 1 || i - 1 + rows > desc[2] || j - 1 + cols > desc[3])
        detail::fail(who + ": submatrix (" + std::to_string(i) + ", " + std::to_string(j) + ") + " +
                     std::to_string(rows) + " x " + std::to_string(cols) + " exceeds the " +
                     std::to_string(desc[2]) + " x " + std::to_string(desc[3]) + " matrix");
    const int lld = scalapack::leading_dimension(desc);
    if (lld < std::max(1, scalapack::numroc(desc[2], desc[4], 0, desc[6], 1)))
        detail::fail(who + ": local leading dimension " + std::to_string(lld) + " < " + std::to_string(desc[2]));
    return (long long)(i - 1) + (long long)(j - 1) * lld;
}

void check_single_process_grid(int ctxt) {
    const scalapack::grid_info g = scalapack::get_grid_info(ctxt);
    if (g.nprow * g.npcol != 1)
        detail::fail("pxgemm: the BLACS grid is " + std::to_string(g.nprow) + " x " + std::to_string(g.npcol) +
                     "; this build runs on a single process (1 x 1 grid)");
}

}  // namespace

template <>
void scale_matrix<double>(const int *descc, double *c, const int ic, const int jc, const int m, const int n,
                          const double beta) {
    if (beta == 1.0 || m == 0 || n == 0) return;
    check_single_process_grid(scalapack::get_grid_context(descc));
    const long long off = submatrix_offset("C", descc, ic, jc, m, n);
    auto ctx = get_context_instance<double>();
    // k == 0: C <- beta * C on the device (beta == 0 writes zeros)
    detail::check(cosma_amd_dgemm_host(ctx->get_gpu_context(), 'N', 'N', m, n, 0, 0.0, nullptr, 1, nullptr, 1, beta,
                                       c + off, scalapack::leading_dimension(descc), 0, 1),
                  "scale_matrix");
}

template <>
void pxgemm<double>(const char trans_a, const char trans_b, const int m, const int n, const int k, const double alpha,
                    const double *a, const int ia, const int ja, const int *desca, const double *b, const int ib,
                    const int jb, const int *descb, const double beta, double *c, const int ic, const int jc,
                    const int *descc) {
    // corner cases allowed by the standard (src/cosma/cosma_pxgemm.cpp:39-46)
    if (m == 0 || n == 0) return;
    if (m < 0 || n < 0 || k < 0) detail::fail("pxgemm: negative dimension");
    if (k == 0 || alpha == 0.0) {
        scale_matrix<double>(descc, c, ic, jc, m, n, beta);
        return;
    }
    const char ta = (char)std::toupper((unsigned char)trans_a), tb = (char)std::toupper((unsigned char)trans_b);
    if (ta != 'N' && ta != 'T' && ta != 'C') detail::fail(std::string("pxgemm: invalid trans_a '") + trans_a + "'");
    if (tb != 'N' && tb != 'T' && tb != 'C') detail::fail(std::string("pxgemm: invalid trans_b '") + trans_b + "'");
    check_single_process_grid(scalapack::get_grid_context(desca, descb, descc));

    auto ctx = get_context_instance<double>();
    // the memory the reference's strategy accounts for (src/cosma/strategy.cpp:373-381)
    const long long need = (long long)m * k + (long long)k * n + (long long)m * n;
    if (ctx->get_cpu_memory_limit() < need)
        detail::fail("This multiplication requires the memory for at least " + std::to_string(need) +
                     " units, but only " + std::to_string(ctx->get_cpu_memory_limit()) + " units are allowed.");

    // stored extents of the submatrices (src/cosma/cosma_pxgemm.cpp:90-97)
    const int a_rows = ta == 'N' ? m : k, a_cols = ta == 'N' ? k : m;
    const int b_rows = tb == 'N' ? k : n, b_cols = tb == 'N' ? n : k;
    const long long oa = submatrix_offset("A", desca, ia, ja, a_rows, a_cols);
    const long long ob = submatrix_offset("B", descb, ib, jb, b_rows, b_cols);
    const long long oc = submatrix_offset("C", descc, ic, jc, m, n);
    detail::check(cosma_amd_dgemm_host(ctx->get_gpu_context(), ta, tb, m, n, k, alpha, a + oa,
                                       scalapack::leading_dimension(desca), b + ob,
                                       scalapack::leading_dimension(descb), beta, c + oc,
                                       scalapack::leading_dimension(descc), ctx->pin_host_buffers ? 1 : 0, 1),
                  "pxgemm");
}

}  // namespace cosma

// ------------------------------------------------ COSMA-prefixed C symbols
namespace {

using dgemm_fn = void (*)(const char *, const char *, const int *, const int *, const int *, const double *,
                          const double *, const int *, const int *, const int *, const double *, const int *,
                          const int *, const int *, const double *, double *, const int *, const int *,
                          const int *);
using sgemm_fn = void (*)(const char *, const char *, const int *, const int *, const int *, const float *,
                          const float *, const int *, const int *, const int *, const float *, const int *,
                          const int *, const int *, const float *, float *, const int *, const int *,
                          const int *);

// ScaLAPACK's own definition of `name` (trailing-underscore spelling), if loaded.
template <typename F>
F scalapack_symbol(const char *name) {
    return reinterpret_cast<F>(cosma::detail::find_symbol(name));
}

void prefixed_pdgemm(const char *ta, const char *tb, const int *m, const int *n, const int *k, const double *alpha,
                     const double *a, const int *ia, const int *ja, const int *desca, const double *b,
                     const int *ib, const int *jb, const int *descb, const double *beta, double *c, const int *ic,
                     const int *jc, const int *descc) {
    if (cosma::is_problem_too_small(*m, *n, *k)) {
        if (auto f = scalapack_symbol<dgemm_fn>("pdgemm_")) {
            f(ta, tb, m, n, k, alpha, a, ia, ja, desca, b, ib, jb, descb, beta, c, ic, jc, descc);
            return;
        }
    }
    cosma::pxgemm<double>(*ta, *tb, *m, *n, *k, *alpha, a, *ia, *ja, desca, b, *ib, *jb, descb, *beta, c, *ic, *jc,
                          descc);
}

template <typename T, typename F>
void forward_or_fail(const char *sym, const char *ta, const char *tb, const int *m, const int *n, const int *k,
                     const T *alpha, const T *a, const int *ia, const int *ja, const int *desca, const T *b,
                     const int *ib, const int *jb, const int *descb, const T *beta, T *c, const int *ic,
                     const int *jc, const int *descc) {
    if (auto f = scalapack_symbol<F>(sym)) {
        f(ta, tb, m, n, k, alpha, a, ia, ja, desca, b, ib, jb, descb, beta, c, ic, jc, descc);
        return;
    }
    cosma::detail::fail(std::string("cosma_amd implements double precision only; no ScaLAPACK ") + sym +
                        " is loaded to take this call");
}

}  // namespace

extern "C" {

#define COSMA_AMD_D_ARGS                                                                                     \
    const char *ta, const char *tb, const int *m, const int *n, const int *k, const double *alpha,           \
        const double *a, const int *ia, const int *ja, const int *desca, const double *b, const int *ib,     \
        const int *jb, const int *descb, const double *beta, double *c, const int *ic, const int *jc,        \
        const int *descc
#define COSMA_AMD_S_ARGS                                                                                     \
    const char *ta, const char *tb, const int *m, const int *n, const int *k, const float *alpha,            \
        const float *a, const int *ia, const int *ja, const int *desca, const float *b, const int *ib,       \
        const int *jb, const int *descb, const float *beta, float *c, const int *ic, const int *jc,          \
        const int *descc
#define COSMA_AMD_PASS ta, tb, m, n, k, alpha, a, ia, ja, desca, b, ib, jb, descb, beta, c, ic, jc, descc

void cosma_pdgemm(COSMA_AMD_D_ARGS) { prefixed_pdgemm(COSMA_AMD_PASS); }
void cosma_pdgemm_(COSMA_AMD_D_ARGS) { prefixed_pdgemm(COSMA_AMD_PASS); }
void COSMA_PDGEMM(COSMA_AMD_D_ARGS) { prefixed_pdgemm(COSMA_AMD_PASS); }
void COSMA_PDGEMM_(COSMA_AMD_D_ARGS) { prefixed_pdgemm(COSMA_AMD_PASS); }

void cosma_psgemm(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("psgemm_", COSMA_AMD_PASS); }
void cosma_psgemm_(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("psgemm_", COSMA_AMD_PASS); }
void COSMA_PSGEMM(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("psgemm_", COSMA_AMD_PASS); }
void COSMA_PSGEMM_(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("psgemm_", COSMA_AMD_PASS); }

void cosma_pcgemm(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("pcgemm_", COSMA_AMD_PASS); }
void cosma_pcgemm_(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("pcgemm_", COSMA_AMD_PASS); }
void COSMA_PCGEMM(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("pcgemm_", COSMA_AMD_PASS); }
void COSMA_PCGEMM_(COSMA_AMD_S_ARGS) { forward_or_fail<float, sgemm_fn>("pcgemm_", COSMA_AMD_PASS); }

void cosma_pzgemm(COSMA_AMD_D_ARGS) { forward_or_fail<double, dgemm_fn>("pzgemm_", COSMA_AMD_PASS); }
void cosma_pzgemm_(COSMA_AMD_D_ARGS) { forward_or_fail<double, dgemm_fn>("pzgemm_", COSMA_AMD_PASS); }
void COSMA_PZGEMM(COSMA_AMD_D_ARGS) { forward_or_fail<double, dgemm_fn>("pzgemm_", COSMA_AMD_PASS); }
void COSMA_PZGEMM_(COSMA_AMD_D_ARGS) { forward_or_fail<double, dgemm_fn>("pzgemm_", COSMA_AMD_PASS); }

}  // extern "C"
