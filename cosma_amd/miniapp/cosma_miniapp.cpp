// cosma_miniapp: C = A * B through cosma::multiply with CosmaMatrix objects,
// timed per repetition (the driver shape of miniapp/cosma_miniapp.cpp:36-83
// in the reference: Strategy from -m -n -k and the -s steps, matrices A, B, C
// on the context's memory pool, alpha = 1, beta = 0, multiply timed with
// steady_clock).
//
// Differences from the reference's miniapp: one rank only; A and B are
// filled with seeded uniform [-1, 1) values (element (i, j) of A is
// u(splitmix64(1 * 0x9E3779B97F4A7C15 + i + j * m)), B likewise with seed 2)
// instead of 10 * drand48(), so a checker can regenerate them; the result
// can be written out (--out) for that checker; timings are printed as one
// JSON line (--json) besides the reference's text.
//
//   cosma_miniapp [-m M] [-n N] [-k K] [-s sm2,sk2] [-r REPS] [--json] [--out FILE]
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <cosma/environment_variables.hpp>
#include <cosma/multiply.hpp>

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

// fills X (in its CosmaMatrix layout) with the global column-major stream of `seed`
void fill(cosma::CosmaMatrix<double> &X, uint64_t seed) {
    double *p = X.matrix_pointer();
    for (const auto &b : X.blocks())
        for (long long c = 0; c < b.cols; ++c)
            for (long long r = 0; r < b.rows; ++r)
                p[b.offset + r + c * b.rows] = uniform_at(seed, (uint64_t)(b.row0 + r) + (uint64_t)(b.col0 + c) * X.m());
}

std::vector<std::string> split_steps(const std::string &s) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) if (!item.empty()) out.push_back(item);
    return out;
}

int usage() {
    std::cerr << "usage: cosma_miniapp [-m M] [-n N] [-k K] [-s steps] [-r reps] [--json] [--out file]\n";
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    int m = 1000, n = 1000, k = 1000, reps = 2;
    std::string steps, out;
    bool json = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        const char *v = nullptr;
        if (a == "-m" && (v = next())) m = std::atoi(v);
        else if (a == "-n" && (v = next())) n = std::atoi(v);
        else if (a == "-k" && (v = next())) k = std::atoi(v);
        else if (a == "-s" && (v = next())) steps = v;
        else if (a == "-r" && (v = next())) reps = std::atoi(v);
        else if (a == "--out" && (v = next())) out = v;
        else if (a == "--json") json = true;
        else return usage();
    }
    if (m <= 0 || n <= 0 || k <= 0 || reps <= 0) return usage();
    try {
        // memory limit from COSMA_CPU_MAX_MEMORY, as the reference's miniapp
        const cosma::Strategy strategy(m, n, k, 1, split_steps(steps), cosma::get_cpu_max_memory<double>());
        if (!json) std::cout << "Strategy = " << strategy << std::endl;
        cosma::CosmaMatrix<double> A('A', strategy, 0), B('B', strategy, 0), C('C', strategy, 0);
        fill(A, 1);
        fill(B, 2);
        std::vector<double> ms;
        for (int r = 0; r < reps; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            cosma::multiply<double>(A, B, C, strategy, 0, 1.0, 0.0);
            const auto t1 = std::chrono::steady_clock::now();
            ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
        }
        if (!out.empty()) {
            std::vector<double> g((size_t)m * n);
            const double *p = C.matrix_pointer();
            for (const auto &b : C.blocks())
                for (long long c = 0; c < b.cols; ++c)
                    for (long long r = 0; r < b.rows; ++r)
                        g[(size_t)(b.row0 + r) + (size_t)(b.col0 + c) * m] = p[b.offset + r + c * b.rows];
            FILE *f = std::fopen(out.c_str(), "wb");
            if (!f || std::fwrite(g.data(), sizeof(double), g.size(), f) != g.size()) {
                std::cerr << "cannot write " << out << "\n";
                return 1;
            }
            std::fclose(f);
        }
        std::vector<double> sorted = ms;
        std::sort(sorted.begin(), sorted.end());
        const double flops = 2.0 * m * n * k;
        if (json) {
            std::printf("{\"m\": %d, \"n\": %d, \"k\": %d, \"steps\": \"%s\", \"reps\": %d, \"ms\": [", m, n, k,
                        steps.c_str(), reps);
            for (size_t i = 0; i < ms.size(); ++i) std::printf("%s%.4f", i ? ", " : "", ms[i]);
            std::printf("], \"min_ms\": %.4f, \"median_ms\": %.4f, \"tflops_min\": %.4f}\n", sorted.front(),
                        sorted[sorted.size() / 2], flops / (sorted.front() * 1e-3) / 1e12);
        } else {
            std::cout << "COSMA TIMES [ms] =";
            for (double t : ms) std::cout << " " << (long)t;
            std::cout << std::endl;
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
