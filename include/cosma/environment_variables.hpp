// Environment knobs of the COSMA interface, read by cosma_amd's front-end.
// Same variable names, defaults and parsing as src/cosma/environment_variables.hpp
// (names and defaults :9-37, parsing src/cosma/environment_variables.cpp), so
// existing job scripts keep their meaning.  What each knob does here:
//   COSMA_GPU_STREAMS             compute streams of the host-streaming GEMM
//   COSMA_GPU_MAX_TILE_{M,N,K}    host-streaming tile sizes
//   COSMA_GPU_MEMORY_PINNING      ON: the memory pool of CosmaMatrix is
//                                 page-locked (hipHostMalloc); OFF: pageable,
//                                 staged through the GEMM's pinned buffers
//                                 (host memory is never registered in place)
//   COSMA_CPU_MAX_MEMORY          memory limit in MB per rank: multiplies that
//                                 need more (m*k + k*n + m*n elements) throw
//   COSMA_DIM_THRESHOLD           pdgemm problems with min(m,n,k) below it are
//                                 passed on to the next pdgemm (ScaLAPACK)
//   COSMA_MEMORY_POOL_AMORTIZATION  growth factor of the memory pool's reserve
//   COSMA_CPU_MEMORY_ALIGNMENT    byte alignment of the pool and its buffers
//   COSMA_ADAPT_STRATEGY, COSMA_OVERLAP_COMM_AND_COMP,
//   COSMA_MIN_LOCAL_DIMENSION, COSMA_GPU_UNIFIED_MEMORY
//                                 read and exposed; they steer multi-rank
//                                 strategies and unified-memory GEMMs that a
//                                 single rank on one MI355X does not have.
#pragma once
#include <cstddef>
#include <limits>
#include <string>

namespace cosma {

namespace env_var_names {
const std::string gpu_n_streams = "COSMA_GPU_STREAMS";
const std::string gpu_tile_m = "COSMA_GPU_MAX_TILE_M";
const std::string gpu_tile_n = "COSMA_GPU_MAX_TILE_N";
const std::string gpu_tile_k = "COSMA_GPU_MAX_TILE_K";
const std::string adapt_strategy = "COSMA_ADAPT_STRATEGY";
const std::string overlap = "COSMA_OVERLAP_COMM_AND_COMP";
const std::string cpu_max_memory = "COSMA_CPU_MAX_MEMORY";
const std::string memory_pinning_enabled = "COSMA_GPU_MEMORY_PINNING";
const std::string memory_pool_amortization = "COSMA_MEMORY_POOL_AMORTIZATION";
const std::string min_local_dimension = "COSMA_MIN_LOCAL_DIMENSION";
const std::string cosma_dim_threshold = "COSMA_DIM_THRESHOLD";
const std::string cosma_cpu_memory_alignment = "COSMA_CPU_MEMORY_ALIGNMENT";
const std::string cosma_gpu_unified_memory = "COSMA_GPU_UNIFIED_MEMORY";
}  // namespace env_var_names

namespace env_var_defaults {
const int gpu_n_streams = 2;
const int gpu_tile_m = 5000;
const int gpu_tile_n = 5000;
const int gpu_tile_k = 5000;
const bool adapt_strategy = true;
const bool overlap = false;
const long long cpu_max_memory = std::numeric_limits<long long>::max();
const bool memory_pinning_enabled = true;
const double memory_pool_amortization = 1.2;
const int min_local_dimension = 200;
const int cosma_dim_threshold = 0;
const int cosma_cpu_memory_alignment = 0;
const bool unified_memory = false;
}  // namespace env_var_defaults

bool env_var_defined(const char *var_name);
bool get_bool_env_var(std::string name, bool default_value);   // "ON" (any case) is true
int get_int_env_var(std::string name, int default_value);
std::size_t get_ull_env_var(std::string name, std::size_t default_value);
float get_float_env_var(std::string name, float default_value);
double get_double_env_var(std::string name, double default_value);

int gpu_streams();
int gpu_max_tile_m();
int gpu_max_tile_n();
int gpu_max_tile_k();
bool get_adapt_strategy();
bool get_overlap_comm_and_comp();
double get_memory_pool_amortization();
// COSMA_CPU_MAX_MEMORY (MB) converted to a number of elements of type T
template <typename T>
long long get_cpu_max_memory();
bool get_memory_pinning();
int get_min_local_dimension();
int get_cosma_dim_threshold();
int get_cosma_cpu_memory_alignment();
bool get_unified_memory();

}  // namespace cosma
