// CosmaMatrix: a matrix in COSMA's native layout, owned by one rank.
//
// Mirrors the user-facing part of src/cosma/matrix.hpp:25-200 (constructors,
// m(), n(), label(), matrix_pointer(), matrix_size(), operator[],
// local_coordinates(), global_coordinates(), rank()).  With one rank the
// local data is the list of leaf blocks of the strategy's sequential steps
// that split this matrix (A: m and k, B: k and n, C: m and n), in depth-first
// order, each block column-major with its own row count as leading
// dimension; with no such steps it is the whole matrix, column-major with
// ld = rows (src/cosma/mapper.cpp:105-200).
//
// The storage is a buffer of the context's memory pool (page-locked host
// memory unless COSMA_GPU_MEMORY_PINNING=OFF), as in the reference
// (src/cosma/buffer.cpp:131-141): its address is looked up in the pool on
// every matrix_pointer(), because the pool may move when it grows.
#pragma once
#include <cstddef>
#include <iosfwd>
#include <memory>
#include <utility>
#include <vector>

#include <cosma/context.hpp>
#include <cosma/strategy.hpp>

namespace cosma {

// Half-open row range [row0, row0 + rows) x column range [col0, col0 + cols)
// stored at offset `offset` of the local buffer with leading dimension rows.
struct LocalBlock {
    long long row0, rows, col0, cols;
    std::size_t offset;
};

template <typename Scalar>
class CosmaMatrix {
  public:
    using scalar_t = Scalar;

    CosmaMatrix(cosma_context<Scalar> *ctxt, char label, const Strategy &strategy, int rank, bool dry_run = false);
    CosmaMatrix(std::unique_ptr<cosma_context<Scalar>> &ctxt, char label, const Strategy &strategy, int rank,
                bool dry_run = false);
    CosmaMatrix(char label, const Strategy &strategy, int rank, bool dry_run = false);
    ~CosmaMatrix();
    CosmaMatrix(const CosmaMatrix &) = delete;
    CosmaMatrix &operator=(const CosmaMatrix &) = delete;

    int m() const { return m_; }
    int n() const { return n_; }
    char label() const { return label_; }
    char which_matrix() const { return label_; }
    int rank() const { return rank_; }

    scalar_t *matrix_pointer();
    const scalar_t *matrix_pointer() const;
    std::size_t matrix_size() const { return size_; }
    std::size_t matrix_size(int rank) const { return rank == 0 ? size_ : 0; }
    scalar_t &operator[](std::size_t index) { return matrix_pointer()[index]; }
    scalar_t operator[](std::size_t index) const { return matrix_pointer()[index]; }

    // (local index, owning rank) of global element (gi, gj)
    std::pair<int, int> local_coordinates(int gi, int gj) const;
    // global (row, column) of a local index
    std::pair<int, int> global_coordinates(int local_index) const;
    std::pair<int, int> global_coordinates(int local_index, int rank) const;

    // leaf blocks of the local buffer, in storage order
    const std::vector<LocalBlock> &blocks() const { return blocks_; }
    // the block covering rows [r0, r0+nr) x columns [c0, c0+nc), or nullptr
    const LocalBlock *find_block(long long r0, long long nr, long long c0, long long nc) const;

    // sizes of the buffers the matrix needs (one: the local matrix)
    std::vector<std::size_t> required_memory() const { return {size_}; }
    // allocates the storage of a matrix constructed with dry_run = true
    void allocate();
    cosma_context<scalar_t> *get_context() { return ctxt_; }

  private:
    void build_layout(const Strategy &strategy);
    cosma_context<scalar_t> *ctxt_;
    char label_;
    int rank_;
    int m_ = 0, n_ = 0;
    std::vector<LocalBlock> blocks_;
    std::size_t size_ = 0;
    bool allocated_ = false;
    std::size_t buffer_id_ = 0;   // offset of the matrix in the context's memory pool
};

template <typename Scalar>
std::ostream &operator<<(std::ostream &os, const CosmaMatrix<Scalar> &mat);

}  // namespace cosma
