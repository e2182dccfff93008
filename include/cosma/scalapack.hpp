// ScaLAPACK descriptor helpers (src/cosma/scalapack.hpp:11-75,
// src/cosma/scalapack.cpp:18-143): descriptor fields, numroc and the
// leading-dimension bounds, with the same meaning.  The BLACS grid of a
// context is asked from the process's BLACS (Cblacs_gridinfo) when one is
// loaded; otherwise every context is the 1 x 1 grid of a single rank.
#pragma once

namespace cosma {
namespace scalapack {

struct block_size {
    int rows = 0, cols = 0;
    block_size() = default;
    block_size(int r, int c) : rows(r), cols(c) {}
    explicit block_size(const int *desc) : rows(desc[4]), cols(desc[5]) {}
};

struct global_matrix_size {
    int rows = 0, cols = 0;
    global_matrix_size() = default;
    global_matrix_size(int r, int c) : rows(r), cols(c) {}
    explicit global_matrix_size(const int *desc) : rows(desc[2]), cols(desc[3]) {}
};

struct rank_src {
    int row_src = 0, col_src = 0;
    rank_src() = default;
    rank_src(int r, int c) : row_src(r), col_src(c) {}
    explicit rank_src(const int *desc) : row_src(desc[6]), col_src(desc[7]) {}
};

struct grid_info {
    int nprow = 1, npcol = 1, myrow = 0, mycol = 0;
};

// BLACS grid of a context (1 x 1 when no BLACS is loaded in the process).
grid_info get_grid_info(int ctxt);

int get_grid_context(const int *desca, const int *descb, const int *descc);
int get_grid_context(const int *desc);
int leading_dimension(const int *desc);

// rows (or columns) of an n-long dimension in blocks of nb that process
// proc_coord owns when the first block is on proc_src, of n_procs
int numroc(int n, int nb, int proc_coord, int proc_src, int n_procs);
int min_leading_dimension(int n, int nb, int rank_grid_dim);
int max_leading_dimension(int n, int nb, int rank_grid_dim);
int local_buffer_size(const int *desc);

}  // namespace scalapack
}  // namespace cosma
