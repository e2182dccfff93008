// Environment knobs (same names, defaults and parsing as
// src/cosma/environment_variables.cpp: "ON" in any case is true for booleans,
// atoi / atof for numbers, COSMA_CPU_MAX_MEMORY in MB converted to elements).
#include <algorithm>
#include <cctype>
#include <complex>
#include <cstdlib>

#include <cosma/environment_variables.hpp>

namespace cosma {

bool env_var_defined(const char *var_name) { return std::getenv(var_name) != nullptr; }

bool get_bool_env_var(std::string name, bool default_value) {
    const char *v = std::getenv(name.c_str());
    if (!v) return default_value;
    std::string s(v);
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return (char)std::toupper(ch); });
    return s == "ON";
}

int get_int_env_var(std::string name, int default_value) {
    const char *v = std::getenv(name.c_str());
    return v ? std::atoi(v) : default_value;
}

std::size_t get_ull_env_var(std::string name, std::size_t default_value) {
    const char *v = std::getenv(name.c_str());
    return v ? (std::size_t)std::stoull(std::string(v)) : default_value;
}

float get_float_env_var(std::string name, float default_value) {
    const char *v = std::getenv(name.c_str());
    return v ? (float)std::atof(v) : default_value;
}

double get_double_env_var(std::string name, double default_value) {
    const char *v = std::getenv(name.c_str());
    return v ? std::atof(v) : default_value;
}

int gpu_streams() { return get_int_env_var(env_var_names::gpu_n_streams, env_var_defaults::gpu_n_streams); }
int gpu_max_tile_m() { return get_int_env_var(env_var_names::gpu_tile_m, env_var_defaults::gpu_tile_m); }
int gpu_max_tile_n() { return get_int_env_var(env_var_names::gpu_tile_n, env_var_defaults::gpu_tile_n); }
int gpu_max_tile_k() { return get_int_env_var(env_var_names::gpu_tile_k, env_var_defaults::gpu_tile_k); }
bool get_adapt_strategy() {
    return get_bool_env_var(env_var_names::adapt_strategy, env_var_defaults::adapt_strategy);
}
bool get_overlap_comm_and_comp() { return get_bool_env_var(env_var_names::overlap, env_var_defaults::overlap); }
bool get_memory_pinning() {
    return get_bool_env_var(env_var_names::memory_pinning_enabled, env_var_defaults::memory_pinning_enabled);
}
bool get_unified_memory() {
    return get_bool_env_var(env_var_names::cosma_gpu_unified_memory, env_var_defaults::unified_memory);
}
double get_memory_pool_amortization() {
    return get_double_env_var(env_var_names::memory_pool_amortization, env_var_defaults::memory_pool_amortization);
}
int get_min_local_dimension() {
    return get_int_env_var(env_var_names::min_local_dimension, env_var_defaults::min_local_dimension);
}
int get_cosma_dim_threshold() {
    return get_int_env_var(env_var_names::cosma_dim_threshold, env_var_defaults::cosma_dim_threshold);
}
int get_cosma_cpu_memory_alignment() {
    return get_int_env_var(env_var_names::cosma_cpu_memory_alignment, env_var_defaults::cosma_cpu_memory_alignment);
}

template <typename T>
long long get_cpu_max_memory() {
    const char *v = std::getenv(env_var_names::cpu_max_memory.c_str());
    if (!v) return env_var_defaults::cpu_max_memory;
    const long long mb = std::atoll(v);
    return mb * 1024LL * 1024LL / (long long)sizeof(T);
}
template long long get_cpu_max_memory<float>();
template long long get_cpu_max_memory<double>();
template long long get_cpu_max_memory<std::complex<float>>();
template long long get_cpu_max_memory<std::complex<double>>();

}  // namespace cosma
