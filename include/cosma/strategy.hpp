// Strategy: the division of an m x n x k multiplication into steps.
//
// Mirrors the constructors and queries of src/cosma/strategy.hpp that user
// code touches (Strategy(m, n, k, P, ...), the step vectors, divisor_* and
// split_* queries, memory accounting; src/cosma/strategy.cpp:82-108,328-381,
// 584-631).  This build runs one rank: P must be 1 and every step must be
// sequential ('s'); a parallel step ('p') or P > 1 throws.  With P = 1 the
// memory needed is m*k + k*n + m*n elements (src/cosma/strategy.cpp:111-117,
// 290-326); a smaller memory_limit throws, as in the reference
// (src/cosma/strategy.cpp:373-381).
#pragma once
#include <cstddef>
#include <iosfwd>
#include <limits>
#include <string>
#include <vector>

namespace cosma {

class Strategy {
  public:
    long long m = 0, n = 0, k = 0;
    std::size_t P = 1;
    long long memory_limit = std::numeric_limits<long long>::max();
    long long memory_used = 0;
    std::vector<int> divisors;       // divisor of each step
    std::string split_dimension;     // 'm', 'n' or 'k' per step
    std::string step_type;           // 's' (sequential) per step
    bool overlap_comm_and_comp = false;

    Strategy() = default;
    Strategy(long long mm, long long nn, long long kk, std::size_t PP,
             long long mem_limit = std::numeric_limits<long long>::max());
    Strategy(long long mm, long long nn, long long kk, std::size_t PP, std::vector<int> &divs,
             std::string &dims, std::string &types,
             long long mem_limit = std::numeric_limits<long long>::max(), bool top_bottom = false);
    // Steps written as "sm2", "sk3", ... (the miniapp's -s syntax).
    Strategy(long long mm, long long nn, long long kk, std::size_t PP, const std::vector<std::string> &steps,
             long long mem_limit = std::numeric_limits<long long>::max());

    std::size_t n_steps() const { return divisors.size(); }
    bool empty() const { return divisors.empty(); }
    bool final_step(std::size_t i) const { return i == n_steps(); }
    bool sequential_step(std::size_t i) const { return step_type[i] == 's'; }
    bool parallel_step(std::size_t i) const { return step_type[i] == 'p'; }
    int divisor(std::size_t i) const { return divisors[i]; }
    bool split_m(std::size_t i) const { return split_dimension[i] == 'm'; }
    bool split_n(std::size_t i) const { return split_dimension[i] == 'n'; }
    bool split_k(std::size_t i) const { return split_dimension[i] == 'k'; }
    bool split_A(std::size_t i) const { return split_m(i) || split_k(i); }
    bool split_B(std::size_t i) const { return split_k(i) || split_n(i); }
    bool split_C(std::size_t i) const { return split_m(i) || split_n(i); }
    bool split(char label, std::size_t i) const;
    int divisor_m(std::size_t i) const { return split_m(i) ? divisors[i] : 1; }
    int divisor_n(std::size_t i) const { return split_n(i) ? divisors[i] : 1; }
    int divisor_k(std::size_t i) const { return split_k(i) ? divisors[i] : 1; }
    int divisor_row(char matrix, std::size_t i) const;
    int divisor_col(char matrix, std::size_t i) const;
    void enable_overlapping_comm_and_comp() { overlap_comm_and_comp = true; }

  private:
    void check_if_valid();
};

std::ostream &operator<<(std::ostream &os, const Strategy &other);

}  // namespace cosma
