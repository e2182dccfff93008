// cosma::pxgemm: C(ic:, jc:) <- alpha * op(A(ia:, ja:)) * op(B(ib:, jb:)) +
// beta * C(ic:, jc:) for matrices in ScaLAPACK block-cyclic layout
// (src/cosma/cosma_pxgemm.hpp:15-35, definition
// src/cosma/cosma_pxgemm.cpp:15-388).
//
// On the 1 x 1 grid of one rank the local array of each matrix is the whole
// global matrix with leading dimension desc[8], so the block-cyclic ->
// COSMA reorder of the reference (a copy, plus a transpose for op = 'T')
// is not needed: op(A) and op(B) are read in place by the transposing DGEMM
// (the host-streaming GEMM when the arrays are in host memory).  Semantics
// kept: m or n == 0 is a no-op; k == 0 or alpha == 0 scales the C submatrix
// by beta (src/cosma/cosma_pxgemm.cpp:39-46, 390-460); beta == 0 never reads C.
// Errors (bad descriptors, a grid with more than one process, too little
// COSMA_CPU_MAX_MEMORY) throw std::runtime_error.
#pragma once
#include <cosma/scalapack.hpp>

namespace cosma {

template <typename T>
void pxgemm(const char trans_a, const char trans_b, const int m, const int n, const int k, const T alpha,
            const T *a, const int ia, const int ja, const int *desca, const T *b, const int ib, const int jb,
            const int *descb, const T beta, T *c, const int ic, const int jc, const int *descc);

// scales the submatrix C(ic:ic+m-1, jc:jc+n-1) by beta (beta == 0 writes zeros)
template <typename T>
void scale_matrix(const int *descc, T *c, const int ic, const int jc, const int m, const int n, const T beta);

// min(m, n, k) < COSMA_DIM_THRESHOLD (src/cosma/cosma_pxgemm.cpp:652-655)
bool is_problem_too_small(int m, int n, int k);

}  // namespace cosma
