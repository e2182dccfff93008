// C API: multiplication of matrices given in arbitrary grid layouts.
//
// Same structs and functions as src/cosma/cinterface.hpp:16-76.  A layout is
// a grid of blocks (row splits, column splits, owner of each block) plus the
// blocks this rank holds; each block is column-major with its own leading
// dimension.  On one rank this rank holds every block.  The blocks are
// gathered into device memory, multiplied once on the MI355X with
// (alpha, beta) = (1, 0), and each block of C is updated as
// C_blk <- alpha * X_blk + beta * C_blk (src/cosma/multiply.cpp:77-213).
// d is implemented; s, c, z report an error.  The communicator must have one
// rank.  Errors throw std::runtime_error (the reference's behaviour).
#pragma once
#include <cosma/mpi_types.hpp>

#ifdef __cplusplus
extern "C" {
#endif

struct block {
    void *data;     // first element of the block
    const int ld;   // leading dimension (distance between columns)
    const int row;  // block row index in the grid
    const int col;  // block column index in the grid
};

struct layout {
    int rowblocks;          // number of block rows of the grid
    int colblocks;          // number of block columns
    const int *rowsplit;    // rows of block row i: [rowsplit[i], rowsplit[i+1])
    const int *colsplit;    // columns of block column j: [colsplit[j], colsplit[j+1])
    const int *owners;      // owner rank of block (i, j) at owners[i * colblocks + j]
    int nlocalblocks;       // blocks held by this rank
    struct block *localblocks;
};

void smultiply_using_layout(MPI_Comm comm, const char *transa, const char *transb, const float *alpha,
                            const struct layout *layout_a, const struct layout *layout_b, const float *beta,
                            const struct layout *layout_c);
void dmultiply_using_layout(MPI_Comm comm, const char *transa, const char *transb, const double *alpha,
                            const struct layout *layout_a, const struct layout *layout_b, const double *beta,
                            const struct layout *layout_c);
void cmultiply_using_layout(MPI_Comm comm, const char *transa, const char *transb, const float *alpha,
                            const struct layout *layout_a, const struct layout *layout_b, const float *beta,
                            const struct layout *layout_c);
void zmultiply_using_layout(MPI_Comm comm, const char *transa, const char *transb, const double *alpha,
                            const struct layout *layout_a, const struct layout *layout_b, const double *beta,
                            const struct layout *layout_c);

#ifdef __cplusplus
}
#endif
