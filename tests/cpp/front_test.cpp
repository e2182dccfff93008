// Driver for the tests of the C++ COSMA interface (tests/test_front*.py).
//
//   front_test strategy
//       Strategy validation and CosmaMatrix layouts (no GPU needed); prints
//       one line per check.
//   front_test local_multiply <variant> m n k alpha beta out.bin
//       cosma::local_multiply on host buffers; variant = singleton | context |
//       handle.  A, B, C are the counter-based uniform matrices with seeds
//       1, 2, 3 (element i of seed s = u(splitmix64(s * 0x9E3779B97F4A7C15 + i))),
//       column-major with ld = m, k, m.  Writes C (m * n doubles).
//   front_test multiply m n k steps alpha beta out.bin
//       cosma::multiply with CosmaMatrix objects and the sequential steps
//       given as "sm2,sk3" ("-" for none).  Matrices hold the same global
//       values as above; writes C in global column-major order.
//   front_test pdgemm ta tb m n k alpha beta ia ja ib jb ic jc out.bin
//       cosma_pdgemm on a 1 x 1 grid: global matrices with 2 extra rows and
//       columns around the submatrices (seeds 1, 2, 3 over the whole local
//       arrays, lld = rows + 3, blocks 32 x 16).  Writes the whole C array.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <cosma/local_multiply.hpp>
#include <cosma/multiply.hpp>
#include <cosma/prefixed_pxgemm.h>

namespace {

uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double uniform_at(uint64_t seed, uint64_t i) {
    return (double)(splitmix64(seed * 0x9E3779B97F4A7C15ULL + i) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

void fill(std::vector<double> &x, uint64_t seed) {
    for (size_t i = 0; i < x.size(); ++i) x[i] = uniform_at(seed, i);
}

void write(const char *path, const double *x, size_t n) {
    FILE *f = std::fopen(path, "wb");
    if (!f || std::fwrite(x, sizeof(double), n, f) != n) throw std::runtime_error("cannot write output");
    std::fclose(f);
}

template <typename F>
void expect_throw(const char *what, F f) {
    try {
        f();
    } catch (const std::runtime_error &e) {
        std::cout << "ok throws " << what << ": " << e.what() << "\n";
        return;
    }
    std::cout << "FAIL no throw " << what << "\n";
}

std::vector<std::string> split_steps(const std::string &s) {
    std::vector<std::string> out;
    if (s == "-") return out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) out.push_back(item);
    return out;
}

int run_strategy() {
    using cosma::Strategy;
    expect_throw("P=2", [] { Strategy s(10, 10, 10, 2); });
    expect_throw("parallel step", [] { Strategy s(10, 10, 10, 1, {"pm2"}); });
    expect_throw("divisor 1", [] { Strategy s(10, 10, 10, 1, {"sm1"}); });
    expect_throw("bad dimension", [] { Strategy s(10, 10, 10, 1, {"sx2"}); });
    expect_throw("divisor above dimension", [] { Strategy s(3, 10, 10, 1, {"sm4"}); });
    expect_throw("memory limit", [] { Strategy s(10, 10, 10, 1, 299); });
    {
        Strategy s(10, 10, 10, 1, 300);
        std::cout << "ok memory_used " << s.memory_used << "\n";
    }
    // layouts: one line per matrix, blocks as row0:rows:col0:cols@offset
    for (const char *steps : {"-", "sm2", "sk3", "sm2,sk2", "sn3,sm2", "sk2,sn2,sm3"}) {
        Strategy s(7, 9, 11, 1, split_steps(steps));
        for (char label : {'A', 'B', 'C'}) {
            cosma::CosmaMatrix<double> X(label, s, 0, true);
            std::cout << "layout " << steps << " " << label << " " << X.m() << "x" << X.n() << " size "
                      << X.matrix_size();
            for (const auto &b : X.blocks())
                std::cout << " " << b.row0 << ":" << b.rows << ":" << b.col0 << ":" << b.cols << "@" << b.offset;
            std::cout << "\n";
        }
    }
    return 0;
}

int run_local_multiply(int argc, char **argv) {
    if (argc != 9) throw std::runtime_error("usage: local_multiply variant m n k alpha beta out");
    const std::string variant = argv[2];
    const int m = std::atoi(argv[3]), n = std::atoi(argv[4]), k = std::atoi(argv[5]);
    const double alpha = std::atof(argv[6]), beta = std::atof(argv[7]);
    std::vector<double> A((size_t)m * k), B((size_t)k * n), C((size_t)m * n);
    fill(A, 1);
    fill(B, 2);
    fill(C, 3);
    if (variant == "singleton") {
        cosma::local_multiply<double>(A.data(), B.data(), C.data(), m, n, k, alpha, beta, true);
    } else if (variant == "context") {
        auto ctx = cosma::make_context<double>(SIZE_MAX, 3, 64, 48, 80);   // small tiles: many streamed tiles
        cosma::local_multiply<double>(ctx, A.data(), B.data(), C.data(), m, n, k, alpha, beta, true);
    } else if (variant == "handle") {
        auto ctx = cosma::make_context<double>();
        cosma::local_multiply<double>(ctx->get_gpu_context(), A.data(), B.data(), C.data(), m, n, k, alpha, beta,
                                      true, true);
    } else {
        throw std::runtime_error("unknown variant " + variant);
    }
    write(argv[8], C.data(), C.size());
    return 0;
}

int run_multiply(int argc, char **argv) {
    if (argc != 9) throw std::runtime_error("usage: multiply m n k steps alpha beta out");
    const int m = std::atoi(argv[2]), n = std::atoi(argv[3]), k = std::atoi(argv[4]);
    const double alpha = std::atof(argv[6]), beta = std::atof(argv[7]);
    cosma::Strategy s(m, n, k, 1, split_steps(argv[5]));
    cosma::CosmaMatrix<double> A('A', s, 0), B('B', s, 0), C('C', s, 0);
    auto fill_global = [](cosma::CosmaMatrix<double> &X, uint64_t seed) {
        for (size_t l = 0; l < X.matrix_size(); ++l) {
            const auto rc = X.global_coordinates((int)l);
            X[l] = uniform_at(seed, (uint64_t)rc.first + (uint64_t)rc.second * X.m());
        }
    };
    fill_global(A, 1);
    fill_global(B, 2);
    fill_global(C, 3);
    cosma::multiply<double>(A, B, C, s, 0, alpha, beta);
    std::vector<double> out((size_t)m * n);
    for (size_t l = 0; l < C.matrix_size(); ++l) {
        const auto rc = C.global_coordinates((int)l);
        out[(size_t)rc.first + (size_t)rc.second * m] = C[l];
    }
    write(argv[8], out.data(), out.size());
    return 0;
}

int run_pdgemm(int argc, char **argv) {
    if (argc != 16) throw std::runtime_error("usage: pdgemm ta tb m n k alpha beta ia ja ib jb ic jc out");
    const char ta = argv[2][0], tb = argv[3][0];
    const int m = std::atoi(argv[4]), n = std::atoi(argv[5]), k = std::atoi(argv[6]);
    const double alpha = std::atof(argv[7]), beta = std::atof(argv[8]);
    const int ia = std::atoi(argv[9]), ja = std::atoi(argv[10]), ib = std::atoi(argv[11]), jb = std::atoi(argv[12]);
    const int ic = std::atoi(argv[13]), jc = std::atoi(argv[14]);
    const bool na = (ta == 'N' || ta == 'n'), nb = (tb == 'N' || tb == 'n');
    // global arrays: the submatrix plus 2 extra rows / columns beyond it
    const int ar = (na ? m : k) + ia + 1, ac = (na ? k : m) + ja + 1;
    const int br = (nb ? k : n) + ib + 1, bc = (nb ? n : k) + jb + 1;
    const int cr = m + ic + 1, cc = n + jc + 1;
    const int lda = ar + 3, ldb = br + 3, ldc = cr + 3;
    std::vector<double> A((size_t)lda * ac), B((size_t)ldb * bc), C((size_t)ldc * cc);
    fill(A, 1);
    fill(B, 2);
    fill(C, 3);
    const int desca[9] = {1, 0, ar, ac, 32, 16, 0, 0, lda};
    const int descb[9] = {1, 0, br, bc, 32, 16, 0, 0, ldb};
    const int descc[9] = {1, 0, cr, cc, 32, 16, 0, 0, ldc};
    cosma_pdgemm(&ta, &tb, &m, &n, &k, &alpha, A.data(), &ia, &ja, desca, B.data(), &ib, &jb, descb, &beta, C.data(),
                 &ic, &jc, descc);
    write(argv[15], C.data(), C.size());
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: front_test strategy|local_multiply|multiply|pdgemm ...\n";
        return 2;
    }
    try {
        const std::string cmd = argv[1];
        if (cmd == "strategy") return run_strategy();
        if (cmd == "local_multiply") return run_local_multiply(argc, argv);
        if (cmd == "multiply") return run_multiply(argc, argv);
        if (cmd == "pdgemm") return run_pdgemm(argc, argv);
        std::cerr << "unknown command " << cmd << "\n";
        return 2;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
