// The local multiply: C <- alpha * A * B + beta * C on column-major buffers
// (A m x k with ld m, B k x n with ld k, C m x n with ld m) that may live in
// host memory (pageable or pinned) or in device memory.
//
// Same overloads as src/cosma/local_multiply.hpp:6-65 (definitions
// src/cosma/local_multiply.cpp:218-381).  Every overload runs on the MI355X:
// the FP64 MFMA DGEMM directly when all three buffers are on the device,
// otherwise the host-streaming GEMM (resident panels or a bounded ring,
// overlapped uploads, per-tile DGEMMs and downloads).  There is no CPU
// branch: local_multiply_cpu, the reference's naive loop that its test
// utilities call for the expected result (utils/cosma_utils.hpp:276-283),
// keeps its signature and contract (host buffers, C <- alpha A B + beta C)
// but runs on the MI355X like the other overloads, so its result equals the
// naive loop's up to FP64 rounding, not bit for bit.  Errors throw
// std::runtime_error.
#pragma once
#include <cosma/context.hpp>

namespace cosma {

template <typename Scalar>
void local_multiply(cosma_context<Scalar> *ctx, Scalar *a, Scalar *b, Scalar *c, int m, int n, int k,
                    Scalar alpha, Scalar beta, bool copy_c_back);

template <typename Scalar>
void local_multiply(context<Scalar> &ctx, Scalar *a, Scalar *b, Scalar *c, int m, int n, int k,
                    Scalar alpha, Scalar beta, bool copy_c_back);

// uses get_context_instance<Scalar>()
template <typename Scalar>
void local_multiply(Scalar *a, Scalar *b, Scalar *c, int m, int n, int k, Scalar alpha, Scalar beta,
                    bool copy_c_back);

// src/cosma/local_multiply.hpp:18-26: host buffers, C overwritten in place
// (uses get_context_instance<Scalar>()).
template <typename Scalar>
void local_multiply_cpu(Scalar *a, Scalar *b, Scalar *c, int m, int n, int k, Scalar alpha, Scalar beta);

// On an explicit device handle (the reference's gpu::mm_handle overload);
// with copy_c_back == false the result stays on the device (see
// cosma_amd_get_device_c).
template <typename Scalar>
void local_multiply(cosma_amd_handle_t gpu_ctx, Scalar *a, Scalar *b, Scalar *c, int m, int n, int k,
                    Scalar alpha, Scalar beta, bool pin_host_buffers, bool copy_c_back);

}  // namespace cosma
