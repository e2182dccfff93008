// Driver for the tests of the C++ COSMA interface (tests/test_front*.py).
//
//   front_test strategy
//       Strategy validation and CosmaMatrix layouts (no GPU needed); prints
//       one line per check.
//   front_test local_multiply <variant> m n k alpha beta out.bin
//       cosma::local_multiply on host buffers; variant = singleton | context |
//       handle | cpu_name (cosma::local_multiply_cpu).  A, B, C are the counter-based uniform matrices with seeds
//       1, 2, 3 (element i of seed s = u(splitmix64(s * 0x9E3779B97F4A7C15 + i))),
//       column-major with ld = m, k, m.  Writes C (m * n doubles).
//   front_test multiply m n k steps alpha beta out.bin
//       cosma::multiply with CosmaMatrix objects and the sequential steps
//       given as "sm2,sk3" ("-" for none).  Matrices hold the same global
//       values as above; writes C in global column-major order.
//   front_test env [handle]
//       The knobs as the interface reads them (one "name value" line each);
//       with "handle" also the streams and tiles of the context's device
//       handle (needs the GPU).
//   front_test pool
//       memory_pool and CosmaMatrix storage checks (no GPU needed with
//       COSMA_GPU_MEMORY_PINNING=OFF); prints "ok ..." / "FAIL ..." lines and
//       the pool's figures.
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

#include <cosma/environment_variables.hpp>
#include <cosma/local_multiply.hpp>
#include <cosma/memory_pool.hpp>
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

int run_env(int argc, char **argv) {
    cosma::cosma_context<double> ctx;
    std::cout << "streams " << ctx.get_gpu_streams() << "\n";
    std::cout << "tiles " << ctx.get_gpu_tile_m() << " " << ctx.get_gpu_tile_n() << " " << ctx.get_gpu_tile_k() << "\n";
    std::cout << "cpu_max_memory " << ctx.get_cpu_memory_limit() << "\n";
    std::cout << "pinning " << ctx.pin_host_buffers << " " << ctx.get_memory_pool().page_locked() << "\n";
    std::cout << "amortization " << ctx.get_memory_pool().amortization << "\n";
    std::cout << "alignment " << cosma::get_cosma_cpu_memory_alignment() << "\n";
    std::cout << "dim_threshold " << cosma::get_cosma_dim_threshold() << "\n";
    std::cout << "adapt " << ctx.adapt_to_scalapack_strategy << "\n";
    std::cout << "overlap " << ctx.overlap_comm_and_comp << "\n";
    std::cout << "unified " << ctx.unified_memory() << "\n";
    std::cout << "min_local_dimension " << cosma::get_min_local_dimension() << "\n";
    if (argc > 2 && std::string(argv[2]) == "handle") {
        int dev = 0, st = 0;
        int64_t tm = 0, tn = 0, tk = 0;
        if (cosma_amd_handle_get_config(ctx.get_gpu_context(), &dev, &st, &tm, &tn, &tk) != COSMA_AMD_OK)
            throw std::runtime_error(cosma_amd_last_error());
        std::cout << "handle " << dev << " " << st << " " << tm << " " << tn << " " << tk << "\n";
    }
    return 0;
}

void expect(bool ok, const std::string &what) { std::cout << (ok ? "ok " : "FAIL ") << what << "\n"; }

int run_pool() {
    // a pool of its own: buffers, growth, order of release
    {
        cosma::memory_pool<double> pool;
        const size_t pad = pool.padded(1001);
        std::vector<size_t> sizes = {1001, 500, 3};
        pool.reserve(sizes);
        std::cout << "reserved " << pool.capacity() << " padded1001 " << pad << " sizes " << sizes[0] << " "
                  << sizes[1] << " " << sizes[2] << "\n";
        const size_t a = pool.get_buffer_id(1001), b = pool.get_buffer_id(500), c = pool.get_buffer_id(3);
        std::cout << "ids " << a << " " << b << " " << c << " size " << pool.size() << "\n";
        double *pa = pool.get_buffer_pointer(a);
        for (int i = 0; i < 1001; ++i) pa[i] = i;
        double *pb = pool.get_buffer_pointer(b);
        for (int i = 0; i < 500; ++i) pb[i] = -i;
        const size_t cap0 = pool.capacity();
        expect(cap0 >= pool.size(), "reserve covers the buffers");
        // a buffer beyond the capacity grows the pool and keeps the contents
        const size_t big = pool.get_buffer_id(10 * cap0);
        double *pbig = pool.get_buffer_pointer(big);
        pbig[0] = 7.0;
        pa = pool.get_buffer_pointer(a);
        pb = pool.get_buffer_pointer(b);
        bool kept = true;
        for (int i = 0; i < 1001; ++i) kept &= pa[i] == i;
        for (int i = 0; i < 500; ++i) kept &= pb[i] == -i;
        expect(pool.capacity() >= pool.size() && pool.capacity() > cap0, "growth beyond the reserve");
        expect(kept, "growth keeps the contents");
        // release out of order: b before the buffers above it
        const size_t top_before = pool.size();
        pool.free_buffer(pool.get_buffer_pointer(b), 500);
        expect(pool.size() == top_before, "a buffer below the top stays until the top is released");
        pool.free_buffer(pool.get_buffer_pointer(big), 10 * cap0);
        expect(pool.size() == c + pool.padded(3), "releasing the top stops at c, still in use");
        pool.free_buffer(pool.get_buffer_pointer(c), 3);
        expect(pool.size() == a + pool.padded(1001), "the released b goes with c");
        pool.free_buffer(pool.get_buffer_pointer(a), 1001);
        expect(pool.size() == 0, "empty");
        std::cout << "pool capacity " << pool.capacity() << " page_locked " << pool.page_locked() << "\n";
    }
    // CosmaMatrix storage comes from the context's pool
    {
        auto ctx = cosma::make_context<double>();
        cosma::Strategy s(30, 20, 40, 1);
        {
            cosma::CosmaMatrix<double> A(ctx, 'A', s, 0), B(ctx, 'B', s, 0);
            for (size_t i = 0; i < A.matrix_size(); ++i) A.matrix_pointer()[i] = (double)i;
            cosma::CosmaMatrix<double> C(ctx, 'C', s, 0);   // may move the pool
            bool ok = true;
            for (size_t i = 0; i < A.matrix_size(); ++i) ok &= A[i] == (double)i;
            expect(ok, "matrix contents survive the pool growing");
            auto &pool = ctx->get_memory_pool();
            expect(pool.size() == pool.padded(1200) + pool.padded(800) + pool.padded(600), "three matrices in the pool");
            expect(A.matrix_pointer() == pool.get_pool_pointer(), "A at the bottom of the pool");
            std::cout << "matrices " << A.matrix_size() << " " << B.matrix_size() << " " << C.matrix_size()
                      << " pool " << pool.size() << "\n";
        }
        expect(ctx->get_memory_pool().size() == 0, "matrices return their buffers");
        expect(ctx->get_cosma_comm() == nullptr, "no communicator on one rank");
        ctx->register_state(0, s);
        expect(!ctx->get_memory_pool().already_pinned, "register_state of a new strategy");
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
    } else if (variant == "cpu_name") {
        cosma::local_multiply_cpu<double>(A.data(), B.data(), C.data(), m, n, k, alpha, beta);
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
        if (cmd == "env") return run_env(argc, argv);
        if (cmd == "pool") return run_pool();
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
