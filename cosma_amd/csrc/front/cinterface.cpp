// C API dmultiply_using_layout on one rank (src/cosma/cinterface.hpp:16-76,
// src/cosma/cinterface.cpp:11-107; the work of cosma::multiply_using_layout,
// src/cosma/multiply.cpp:77-213).  The layouts are validated and turned into
// block lists for cosma_amd_dgemm_layout, which gathers them on the device,
// multiplies once and applies alpha and beta per block of C.
#include <cctype>
#include <string>
#include <vector>

#include <cosma/cinterface.hpp>
#include <cosma/context.hpp>

#include "front.hpp"

namespace {

struct Extent {
    int64_t rows, cols;
};

Extent blocks_of(const char *name, const layout *l, std::vector<cosma_amd_block> &out) {
    const std::string who = std::string("multiply_using_layout: layout of ") + name;
    if (!l) cosma::detail::fail(who + " is null");
    if (l->rowblocks < 1 || l->colblocks < 1 || !l->rowsplit || !l->colsplit)
        cosma::detail::fail(who + " has no grid");
    if (l->rowsplit[0] != 0 || l->colsplit[0] != 0) cosma::detail::fail(who + ": splits must start at 0");
    for (int i = 0; i < l->rowblocks; ++i)
        if (l->rowsplit[i + 1] < l->rowsplit[i]) cosma::detail::fail(who + ": row splits decrease");
    for (int j = 0; j < l->colblocks; ++j)
        if (l->colsplit[j + 1] < l->colsplit[j]) cosma::detail::fail(who + ": column splits decrease");
    if (l->owners)
        for (int b = 0; b < l->rowblocks * l->colblocks; ++b)
            if (l->owners[b] != 0)
                cosma::detail::fail(who + ": block " + std::to_string(b) + " is owned by rank " +
                                    std::to_string(l->owners[b]) + "; this build runs one rank");
    if (l->nlocalblocks != l->rowblocks * l->colblocks)
        cosma::detail::fail(who + ": " + std::to_string(l->nlocalblocks) + " local blocks, the grid has " +
                            std::to_string(l->rowblocks * l->colblocks) + " (one rank holds them all)");
    for (int b = 0; b < l->nlocalblocks; ++b) {
        const block &blk = l->localblocks[b];
        if (blk.row < 0 || blk.row >= l->rowblocks || blk.col < 0 || blk.col >= l->colblocks)
            cosma::detail::fail(who + ": local block " + std::to_string(b) + " has grid index out of range");
        cosma_amd_block d;
        d.data = blk.data;
        d.ld = blk.ld;
        d.row = l->rowsplit[blk.row];
        d.col = l->colsplit[blk.col];
        d.rows = l->rowsplit[blk.row + 1] - d.row;
        d.cols = l->colsplit[blk.col + 1] - d.col;
        out.push_back(d);
    }
    return {l->rowsplit[l->rowblocks], l->colsplit[l->colblocks]};
}

}  // namespace

extern "C" {

void dmultiply_using_layout(MPI_Comm comm, const char *transa, const char *transb, const double *alpha,
                            const layout *layout_a, const layout *layout_b, const double *beta,
                            const layout *layout_c) {
    const int P = cosma::detail::comm_size(comm);
    if (P != 1)
        cosma::detail::fail("multiply_using_layout: the communicator has " + std::to_string(P) +
                            " ranks; this build runs one");
    const char ta = (char)std::toupper((unsigned char)*transa), tb = (char)std::toupper((unsigned char)*transb);
    std::vector<cosma_amd_block> A, B, C;
    const Extent ea = blocks_of("A", layout_a, A), eb = blocks_of("B", layout_b, B), ec = blocks_of("C", layout_c, C);
    const int64_t m = ec.rows, n = ec.cols;
    const int64_t k = ta == 'N' ? ea.cols : ea.rows;
    if ((ta == 'N' ? ea.rows : ea.cols) != m || (tb == 'N' ? eb.cols : eb.rows) != n ||
        (tb == 'N' ? eb.rows : eb.cols) != k)
        cosma::detail::fail("multiply_using_layout: the global sizes of op(A), op(B) and C do not match");
    auto ctx = cosma::get_context_instance<double>();
    cosma::detail::check(cosma_amd_dgemm_layout(ctx->get_gpu_context(), ta, tb, m, n, k, *alpha, A.data(),
                                                (int64_t)A.size(), B.data(), (int64_t)B.size(), *beta, C.data(),
                                                (int64_t)C.size()),
                         "multiply_using_layout");
}

static void only_double(const char *name) {
    cosma::detail::fail(std::string(name) + ": cosma_amd implements double precision only");
}

void smultiply_using_layout(MPI_Comm, const char *, const char *, const float *, const layout *, const layout *,
                            const float *, const layout *) {
    only_double("smultiply_using_layout");
}

void cmultiply_using_layout(MPI_Comm, const char *, const char *, const float *, const layout *, const layout *,
                            const float *, const layout *) {
    only_double("cmultiply_using_layout");
}

void zmultiply_using_layout(MPI_Comm, const char *, const char *, const double *, const layout *, const layout *,
                            const double *, const layout *) {
    only_double("zmultiply_using_layout");
}

}  // extern "C"
