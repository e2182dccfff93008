// Driver of the host-only sanitizer runs (tests/test_sanitizers.py).
//
//   sanitize_test host_copy   the threaded 2D packing of the host-streaming
//                             GEMM on blocks that use 1 to 8 threads,
//                             sub-blocks with leading dimensions > rows,
//                             and empty blocks; every element compared
//   sanitize_test oracle      the CPU oracle (oracle/oracle.c): the naive
//                             loop, dgemm and double-double dgemm on every
//                             transpose, the block transform, numroc and the
//                             block-cyclic gather, checked for consistency
// Prints "ok <what>" / "FAIL <what>" lines; exit status 1 on any failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host/host_copy.hpp"

extern "C" {
typedef int64_t i64;
void oracle_fill_uniform(double *x, i64 n, uint64_t seed, uint64_t offset);
void oracle_local_multiply_cpu(double *a, double *b, double *c, int m, int n, int k, double alpha, double beta);
int oracle_dgemm(char transa, char transb, i64 m, i64 n, i64 k, double alpha, const double *a, i64 lda,
                 const double *b, i64 ldb, double beta, double *c, i64 ldc);
int oracle_dgemm_dd(char transa, char transb, i64 m, i64 n, i64 k, double alpha, const double *a, i64 lda,
                    const double *b, i64 ldb, double beta, double *c, i64 ldc);
int oracle_block_transform(char op, i64 rows, i64 cols, double alpha, const double *src, i64 lds, double beta,
                           double *dst, i64 ldd);
i64 oracle_numroc(i64 n, i64 nb, int proc_coord, int proc_src, int n_procs);
void oracle_block_cyclic_gather(const double *local, i64 lld, i64 mb, i64 nb, int rsrc, int csrc, int nprow,
                                int npcol, int myrow, int mycol, i64 ia, i64 ja, i64 m, i64 n, double *dst,
                                i64 ldd);
}

namespace {

int failures = 0;

void expect(bool ok, const std::string &what) {
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what.c_str());
    if (!ok) ++failures;
}

std::vector<double> uniform(i64 n, uint64_t seed) {
    std::vector<double> x((size_t)n);
    oracle_fill_uniform(x.data(), n, seed, 0);
    return x;
}

void run_host_copy() {
    struct Case { int64_t rows, cols, sld, dld; };
    // 4 MB per thread: 1 thread, several, the cap of 8, and leading dimensions above rows
    const Case cases[] = {{1, 1, 1, 1}, {100, 37, 100, 103}, {1024, 700, 1030, 1024}, {4096, 1100, 4100, 4096},
                          {513, 2049, 600, 513}, {7, 0, 7, 7}, {0, 9, 1, 1}};
    for (const Case &c : cases) {
        std::vector<double> src((size_t)std::max<int64_t>(1, c.sld * std::max<int64_t>(c.cols, 1)));
        for (size_t i = 0; i < src.size(); ++i) src[i] = (double)i;
        std::vector<double> dst((size_t)std::max<int64_t>(1, c.dld * std::max<int64_t>(c.cols, 1)), -1.0);
        cosma_amd::host_copy_2d(dst.data(), c.dld, src.data(), c.sld, c.rows, c.cols);
        bool ok = true;
        for (int64_t j = 0; j < c.cols; ++j)
            for (int64_t i = 0; i < c.dld; ++i) {
                const double want = i < c.rows ? src[(size_t)(i + j * c.sld)] : -1.0;
                ok &= dst[(size_t)(i + j * c.dld)] == want;
            }
        expect(ok, "host_copy_2d " + std::to_string(c.rows) + "x" + std::to_string(c.cols));
    }
}

double max_rel(const std::vector<double> &x, const std::vector<double> &y) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
        num = std::max(num, std::fabs(x[i] - y[i]));
        den = std::max(den, std::fabs(y[i]));
    }
    return den > 0 ? num / den : num;
}

void run_oracle() {
    const int m = 23, n = 17, k = 31;
    for (const char *t : {"NN", "NT", "TN", "TT"}) {
        const char ta = t[0], tb = t[1];
        const i64 ar = ta == 'N' ? m : k, ac = ta == 'N' ? k : m, br = tb == 'N' ? k : n, bc = tb == 'N' ? n : k;
        std::vector<double> A = uniform(ar * ac, 1), B = uniform(br * bc, 2), C0 = uniform((i64)m * n, 3);
        std::vector<double> c1 = C0, c2 = C0;
        oracle_dgemm(ta, tb, m, n, k, -1.25, A.data(), ar, B.data(), br, 0.5, c1.data(), m);
        oracle_dgemm_dd(ta, tb, m, n, k, -1.25, A.data(), ar, B.data(), br, 0.5, c2.data(), m);
        expect(max_rel(c1, c2) < 1e-13, std::string("dgemm vs double-double ") + t);
        if (ta == 'N' && tb == 'N') {
            std::vector<double> c3 = C0;
            oracle_local_multiply_cpu(A.data(), B.data(), c3.data(), m, n, k, -1.25, 0.5);
            expect(max_rel(c3, c2) < 1e-13, "naive loop vs double-double");
        }
    }
    // block transform: transpose with alpha = 1, beta = 0 is a permutation
    std::vector<double> src = uniform(40 * 30, 4), dst((size_t)33 * 40, 0.0);
    oracle_block_transform('T', 30, 40, 1.0, src.data(), 40, 0.0, dst.data(), 33);
    bool ok = true;
    for (int j = 0; j < 40; ++j)
        for (int i = 0; i < 30; ++i) ok &= dst[(size_t)(i + j * 33)] == src[(size_t)(j + i * 40)];
    expect(ok, "block transform 'T' exact");
    // numroc over a 3-process dimension covers every index once
    i64 total = 0;
    for (int p = 0; p < 3; ++p) total += oracle_numroc(100, 7, p, 1, 3);
    expect(total == 100, "numroc sums to n");
    // gather of a 1 x 1 grid is the identity on the submatrix
    std::vector<double> local = uniform(50 * 40, 5), out((size_t)20 * 10, 0.0);
    oracle_block_cyclic_gather(local.data(), 50, 8, 8, 0, 0, 1, 1, 0, 0, 3, 5, 20, 10, out.data(), 20);
    ok = true;
    for (int j = 0; j < 10; ++j)
        for (int i = 0; i < 20; ++i) ok &= out[(size_t)(i + j * 20)] == local[(size_t)(3 + i + (5 + j) * 50)];   // (ia, ja) 0-based
    expect(ok, "block-cyclic gather on a 1 x 1 grid");
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: sanitize_test host_copy|oracle\n");
        return 2;
    }
    const std::string mode = argv[1];
    if (mode == "host_copy") run_host_copy();
    else if (mode == "oracle") run_oracle();
    else return 2;
    return failures ? 1 : 0;
}
