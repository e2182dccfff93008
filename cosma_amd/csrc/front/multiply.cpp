// Strategy, CosmaMatrix and cosma::multiply for one rank.
//
// Strategy follows the validation rules of src/cosma/strategy.cpp:644-760
// (divisors > 1, dimensions m/n/k, step types p/s) restricted to what a single
// rank can run: P = 1 and sequential steps only.  Memory accounting for P = 1
// (src/cosma/strategy.cpp:111-117, 290-381): m*k + k*n + m*n elements.
//
// CosmaMatrix lays its local data out as the reference's Mapper does for one
// rank (src/cosma/mapper.cpp:105-200): the leaves of the sequential steps
// that split the matrix, depth first, each block column-major with ld = its
// rows.  Intervals split as [len*i/d, len*(i+1)/d) (src/cosma/interval.cpp:84-96).
//
// multiply walks the same steps (src/cosma/multiply.cpp:460-588): splits of m
// and n visit disjoint sub-problems, splits of k reuse the C block with beta
// for the first chunk and 1 afterwards; each leaf is one local multiply.
#include <algorithm>
#include <iostream>
#include <ostream>
#include <sstream>

#include <cosma/local_multiply.hpp>
#include <cosma/multiply.hpp>

#include "front.hpp"

namespace cosma {

// ------------------------------------------------------------- Strategy

Strategy::Strategy(long long mm, long long nn, long long kk, std::size_t PP, long long mem_limit)
    : m(mm), n(nn), k(kk), P(PP), memory_limit(mem_limit) {
    check_if_valid();
}

Strategy::Strategy(long long mm, long long nn, long long kk, std::size_t PP, std::vector<int> &divs,
                   std::string &dims, std::string &types, long long mem_limit, bool)
    : m(mm), n(nn), k(kk), P(PP), memory_limit(mem_limit), divisors(divs), split_dimension(dims),
      step_type(types) {
    check_if_valid();
}

Strategy::Strategy(long long mm, long long nn, long long kk, std::size_t PP, const std::vector<std::string> &steps,
                   long long mem_limit)
    : m(mm), n(nn), k(kk), P(PP), memory_limit(mem_limit) {
    for (const std::string &s : steps) {
        if (s.size() < 3) detail::fail("Strategy: step '" + s + "' is not of the form <p|s><m|n|k><divisor>");
        step_type.push_back(s[0]);
        split_dimension.push_back(s[1]);
        divisors.push_back(std::atoi(s.c_str() + 2));
    }
    check_if_valid();
}

bool Strategy::split(char label, std::size_t i) const {
    return label == 'A' ? split_A(i) : label == 'B' ? split_B(i) : split_C(i);
}

int Strategy::divisor_row(char matrix, std::size_t i) const {
    return matrix == 'A' ? divisor_m(i) : matrix == 'B' ? divisor_k(i) : matrix == 'C' ? divisor_m(i) : 1;
}

int Strategy::divisor_col(char matrix, std::size_t i) const {
    return matrix == 'A' ? divisor_k(i) : matrix == 'B' ? divisor_n(i) : matrix == 'C' ? divisor_n(i) : 1;
}

void Strategy::check_if_valid() {
    if (m < 0 || n < 0 || k < 0) detail::fail("Strategy: negative dimension");
    if (P != 1)
        detail::fail("Strategy: P = " + std::to_string(P) + " ranks; this build runs one rank (P = 1)");
    if (divisors.size() != split_dimension.size() || divisors.size() != step_type.size())
        detail::fail("Strategy: divisors, split dimensions and step types differ in length");
    long long mi = m, ni = n, ki = k;
    for (std::size_t i = 0; i < divisors.size(); ++i) {
        if (divisors[i] <= 1)
            detail::fail("Strategy: divisors in each step must be larger than 1 (step " + std::to_string(i) +
                         " has " + std::to_string(divisors[i]) + ")");
        const char d = split_dimension[i];
        if (d != 'm' && d != 'n' && d != 'k') detail::fail("Strategy: split dimension must be m, n or k");
        if (step_type[i] == 'p') detail::fail("Strategy: parallel steps need more than one rank");
        if (step_type[i] != 's') detail::fail("Strategy: step type must be p or s");
        long long &len = d == 'm' ? mi : d == 'n' ? ni : ki;
        if (len < divisors[i])
            detail::fail(std::string("Strategy: dimension ") + d + " (" + std::to_string(len) +
                         ") is smaller than the divisor of step " + std::to_string(i));
        len /= divisors[i];
    }
    // one rank holds A, B and C whole (src/cosma/strategy.cpp:111-117)
    memory_used = m * k + k * n + m * n;
    if (memory_limit < memory_used)
        detail::fail("This multiplication requires the memory for at least " + std::to_string(memory_used) +
                     " units, but only " + std::to_string(memory_limit) +
                     " units are allowed. Either increase the memory limit or change the strategy by using "
                     "more sequential steps.");
}

std::ostream &operator<<(std::ostream &os, const Strategy &s) {
    os << "Matrix dimensions (m, n, k) = (" << s.m << ", " << s.n << ", " << s.k << ")\n";
    os << "Number of processors: " << s.P << "\n";
    os << "Overlap of communication and computation: " << (s.overlap_comm_and_comp ? "ON" : "OFF") << ".\n";
    os << "Divisions strategy: \n";
    for (std::size_t i = 0; i < s.n_steps(); ++i)
        os << (s.step_type[i] == 'p' ? "parallel (" : "sequential (") << s.split_dimension[i] << " / "
           << s.divisors[i] << ")\n";
    os << "Required memory per rank (in #elements): " << s.memory_used << "\n";
    os << "Available memory per rank (in #elements): ";
    if (s.memory_limit < std::numeric_limits<long long>::max())
        os << s.memory_limit;
    else
        os << "not specified (assumed: infinite)";
    return os << "\n";
}

// ------------------------------------------------------------ intervals

namespace {

struct Range {   // [first, first + len)
    long long first, len;
    // part i of d: [len*i/d, len*(i+1)/d) (src/cosma/interval.cpp:84-96)
    Range part(int d, int i) const {
        if (len < d) return *this;
        return {first + len * i / d, len * (i + 1) / d - len * i / d};
    }
};

// Rows of the matrix are split by the steps' divisor_row, columns by
// divisor_col; a sequential step over a dimension the matrix does not have
// visits its sub-problem once.
void leaves(const Strategy &s, char label, std::size_t step, Range r, Range c, std::vector<LocalBlock> &out,
            std::size_t &offset) {
    if (s.final_step(step) || s.empty()) {
        out.push_back({r.first, r.len, c.first, c.len, offset});
        offset += (std::size_t)(r.len * c.len);
        return;
    }
    const int dr = s.divisor_row(label, step), dc = s.divisor_col(label, step);
    const int d = s.divisor(step);
    for (int i = 0; i < d; ++i) {
        leaves(s, label, step + 1, r.part(dr, dr > 1 ? i : 0), c.part(dc, dc > 1 ? i : 0), out, offset);
        if (dr * dc == 1) break;
    }
}

}  // namespace

// ----------------------------------------------------------- CosmaMatrix

template <typename Scalar>
CosmaMatrix<Scalar>::CosmaMatrix(cosma_context<Scalar> *ctxt, char label, const Strategy &strategy, int rank,
                                 bool dry_run)
    : ctxt_(ctxt), label_((char)std::toupper((unsigned char)label)), rank_(rank) {
    if (label_ != 'A' && label_ != 'B' && label_ != 'C') detail::fail("CosmaMatrix: label must be A, B or C");
    if (rank != 0) detail::fail("CosmaMatrix: rank " + std::to_string(rank) + " on a single-rank strategy");
    m_ = (int)(label_ == 'A' ? strategy.m : label_ == 'B' ? strategy.k : strategy.m);
    n_ = (int)(label_ == 'A' ? strategy.k : label_ == 'B' ? strategy.n : strategy.n);
    build_layout(strategy);
    if (!dry_run) allocate();
}

template <typename Scalar>
CosmaMatrix<Scalar>::CosmaMatrix(std::unique_ptr<cosma_context<Scalar>> &ctxt, char label, const Strategy &strategy,
                                 int rank, bool dry_run)
    : CosmaMatrix(ctxt.get(), label, strategy, rank, dry_run) {}

template <typename Scalar>
CosmaMatrix<Scalar>::CosmaMatrix(char label, const Strategy &strategy, int rank, bool dry_run)
    : CosmaMatrix(get_context_instance<Scalar>(), label, strategy, rank, dry_run) {}

template <typename Scalar>
CosmaMatrix<Scalar>::~CosmaMatrix() {
    if (allocated_) {
        auto &pool = ctxt_->get_memory_pool();
        pool.free_buffer(pool.get_buffer_pointer(buffer_id_), size_);
    }
}

template <typename Scalar>
Scalar *CosmaMatrix<Scalar>::matrix_pointer() {
    return allocated_ ? ctxt_->get_memory_pool().get_buffer_pointer(buffer_id_) : nullptr;
}

template <typename Scalar>
const Scalar *CosmaMatrix<Scalar>::matrix_pointer() const {
    return allocated_ ? ctxt_->get_memory_pool().get_buffer_pointer(buffer_id_) : nullptr;
}

template <typename Scalar>
void CosmaMatrix<Scalar>::build_layout(const Strategy &strategy) {
    blocks_.clear();
    std::size_t offset = 0;
    leaves(strategy, label_, 0, Range{0, m_}, Range{0, n_}, blocks_, offset);
    size_ = offset;
}

template <typename Scalar>
void CosmaMatrix<Scalar>::allocate() {
    if (allocated_ || size_ == 0) return;
    auto &pool = ctxt_->get_memory_pool();
    buffer_id_ = pool.get_buffer_id(size_);
    allocated_ = true;
    Scalar *data = pool.get_buffer_pointer(buffer_id_);
    std::fill(data, data + size_, Scalar{0});
}

template <typename Scalar>
const LocalBlock *CosmaMatrix<Scalar>::find_block(long long r0, long long nr, long long c0, long long nc) const {
    for (const LocalBlock &b : blocks_)
        if (b.row0 == r0 && b.rows == nr && b.col0 == c0 && b.cols == nc) return &b;
    return nullptr;
}

template <typename Scalar>
std::pair<int, int> CosmaMatrix<Scalar>::local_coordinates(int gi, int gj) const {
    for (const LocalBlock &b : blocks_)
        if (gi >= b.row0 && gi < b.row0 + b.rows && gj >= b.col0 && gj < b.col0 + b.cols)
            return {(int)(b.offset + (gi - b.row0) + (gj - b.col0) * b.rows), 0};
    return {-1, -1};
}

template <typename Scalar>
std::pair<int, int> CosmaMatrix<Scalar>::global_coordinates(int local_index) const {
    for (const LocalBlock &b : blocks_) {
        const long long rel = (long long)local_index - (long long)b.offset;
        if (rel >= 0 && rel < b.rows * b.cols) return {(int)(b.row0 + rel % b.rows), (int)(b.col0 + rel / b.rows)};
    }
    return {-1, -1};
}

template <typename Scalar>
std::pair<int, int> CosmaMatrix<Scalar>::global_coordinates(int local_index, int rank) const {
    return rank == 0 ? global_coordinates(local_index) : std::pair<int, int>{-1, -1};
}

template <typename Scalar>
std::ostream &operator<<(std::ostream &os, const CosmaMatrix<Scalar> &mat) {
    for (std::size_t local = 0; local < mat.matrix_size(); ++local) {
        const auto rc = mat.global_coordinates((int)local);
        os << rc.first << " " << rc.second << " " << mat[local] << std::endl;
    }
    return os;
}

template class CosmaMatrix<double>;
template std::ostream &operator<<(std::ostream &, const CosmaMatrix<double> &);

// -------------------------------------------------------------- multiply

namespace {

template <typename Scalar>
struct Walk {
    cosma_context<Scalar> *ctx;
    CosmaMatrix<Scalar> &A, &B, &C;
    const Strategy &s;
    Scalar alpha;

    Scalar *block(CosmaMatrix<Scalar> &X, Range r, Range c) {
        const LocalBlock *b = X.find_block(r.first, r.len, c.first, c.len);
        if (!b) detail::fail(std::string("multiply: matrix ") + X.label() + " is not in the strategy's layout");
        return X.matrix_pointer() + b->offset;
    }

    void run(std::size_t step, Range m, Range n, Range k, Scalar beta) {
        if (s.final_step(step)) {
            local_multiply<Scalar>(ctx, block(A, m, k), block(B, k, n), block(C, m, n), (int)m.len, (int)n.len,
                                   (int)k.len, alpha, beta, true);
            return;
        }
        const int d = s.divisor(step);
        for (int i = 0; i < d; ++i) {
            if (s.split_m(step)) run(step + 1, m.part(d, i), n, k, beta);
            else if (s.split_n(step)) run(step + 1, m, n.part(d, i), k, beta);
            else run(step + 1, m, n, k.part(d, i), i == 0 ? beta : Scalar{1});
        }
    }
};

}  // namespace

template <typename Scalar>
void multiply(cosma_context<Scalar> *ctx, CosmaMatrix<Scalar> &A, CosmaMatrix<Scalar> &B, CosmaMatrix<Scalar> &C,
              const Strategy &strategy, MPI_Comm comm, Scalar alpha, Scalar beta) {
    if (strategy.m == 0 || strategy.n == 0 || strategy.k == 0) return;   // src/cosma/multiply.cpp:251-254
    const int P = detail::comm_size(comm);
    if (P != 1) detail::fail("multiply: the communicator has " + std::to_string(P) + " ranks; this build runs one");
    if (A.m() != strategy.m || A.n() != strategy.k || B.m() != strategy.k || B.n() != strategy.n ||
        C.m() != strategy.m || C.n() != strategy.n)
        detail::fail("multiply: matrix shapes do not match the strategy");
    if (!A.matrix_pointer() || !B.matrix_pointer() || !C.matrix_pointer())
        detail::fail("multiply: a matrix has no storage (constructed with dry_run and not allocated)");
    ctx->register_state(comm, strategy);
    Walk<Scalar> w{ctx, A, B, C, strategy, alpha};
    w.run(0, Range{0, strategy.m}, Range{0, strategy.n}, Range{0, strategy.k}, beta);
}

template <typename Scalar>
void multiply(CosmaMatrix<Scalar> &A, CosmaMatrix<Scalar> &B, CosmaMatrix<Scalar> &C, const Strategy &strategy,
              MPI_Comm comm, Scalar alpha, Scalar beta) {
    multiply<Scalar>(get_context_instance<Scalar>(), A, B, C, strategy, comm, alpha, beta);
}

template void multiply<double>(cosma_context<double> *, CosmaMatrix<double> &, CosmaMatrix<double> &,
                               CosmaMatrix<double> &, const Strategy &, MPI_Comm, double, double);
template void multiply<double>(CosmaMatrix<double> &, CosmaMatrix<double> &, CosmaMatrix<double> &,
                               const Strategy &, MPI_Comm, double, double);

}  // namespace cosma
