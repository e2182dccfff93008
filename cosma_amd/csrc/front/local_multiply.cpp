// cosma::local_multiply on the MI355X (overloads of
// src/cosma/local_multiply.hpp:6-65; the reference's GPU branch pins the
// buffers and calls Tiled-MM, src/cosma/local_multiply.cpp:218-269,299-381).
// Here every overload is one cosma_amd_dgemm_host call: column-major NN with
// lda = m, ldb = k, ldc = m, operands in host (pageable or pinned) or device
// memory.
#include <cosma/local_multiply.hpp>

#include "front.hpp"

namespace cosma {

template <>
void local_multiply<double>(cosma_amd_handle_t gpu_ctx, double *a, double *b, double *c, int m, int n, int k,
                            double alpha, double beta, bool pin_host_buffers, bool copy_c_back) {
    if (m < 0 || n < 0 || k < 0) detail::fail("local_multiply: negative dimension");
    detail::check(cosma_amd_dgemm_host(gpu_ctx, 'N', 'N', m, n, k, alpha, a, m > 0 ? m : 1, b, k > 0 ? k : 1, beta,
                                       c, m > 0 ? m : 1, pin_host_buffers ? 1 : 0, copy_c_back ? 1 : 0),
                  "local_multiply");
}

template <>
void local_multiply<double>(cosma_context<double> *ctx, double *a, double *b, double *c, int m, int n, int k,
                            double alpha, double beta, bool copy_c_back) {
    local_multiply<double>(ctx->get_gpu_context(), a, b, c, m, n, k, alpha, beta, ctx->pin_host_buffers,
                           copy_c_back);
}

template <>
void local_multiply<double>(context<double> &ctx, double *a, double *b, double *c, int m, int n, int k,
                            double alpha, double beta, bool copy_c_back) {
    local_multiply<double>(ctx.get(), a, b, c, m, n, k, alpha, beta, copy_c_back);
}

template <>
void local_multiply<double>(double *a, double *b, double *c, int m, int n, int k, double alpha, double beta,
                            bool copy_c_back) {
    local_multiply<double>(get_context_instance<double>(), a, b, c, m, n, k, alpha, beta, copy_c_back);
}

// The reference's naive host loop (src/cosma/local_multiply.cpp:277-297):
// same arguments and result, computed by the MI355X DGEMM with C copied back.
template <>
void local_multiply_cpu<double>(double *a, double *b, double *c, int m, int n, int k, double alpha, double beta) {
    local_multiply<double>(get_context_instance<double>(), a, b, c, m, n, k, alpha, beta, true);
}

}  // namespace cosma
