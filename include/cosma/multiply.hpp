// cosma::multiply on one rank: C <- alpha * A * B + beta * C for matrices in
// COSMA's native layout (src/cosma/multiply.hpp:48-54, definition
// src/cosma/multiply.cpp:242-314).  Every leaf problem of the strategy's
// sequential steps is one local multiply on the MI355X; splits of k
// accumulate with beta = 1 after the first chunk (src/cosma/multiply.cpp:
// 460-588).  The communicator must have one rank (checked when MPI is
// loaded in the process).
#pragma once
#include <cosma/context.hpp>
#include <cosma/matrix.hpp>
#include <cosma/mpi_types.hpp>
#include <cosma/strategy.hpp>

namespace cosma {

template <typename Scalar>
void multiply(CosmaMatrix<Scalar> &A, CosmaMatrix<Scalar> &B, CosmaMatrix<Scalar> &C, const Strategy &strategy,
              MPI_Comm comm, Scalar alpha, Scalar beta);

template <typename Scalar>
void multiply(cosma_context<Scalar> *ctx, CosmaMatrix<Scalar> &A, CosmaMatrix<Scalar> &B, CosmaMatrix<Scalar> &C,
              const Strategy &strategy, MPI_Comm comm, Scalar alpha, Scalar beta);

}  // namespace cosma
