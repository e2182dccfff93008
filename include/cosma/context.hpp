// cosma_context: the per-process state of the COSMA interface on one MI355X.
//
// Mirrors the interface of src/cosma/context.hpp:19-66 (constructors
// src/cosma/context.cpp:15-52, singleton :154-158): a context owns the GPU
// handle that the local multiplies run on and the flags read from the
// environment.  Here the GPU handle is a cosma_amd handle (its compute,
// host->device and device->host streams, pinned staging ring and device
// workspaces) instead of a Tiled-MM mm_handle, and there is no communicator
// state: the build covers one rank.
#pragma once
#include <cstddef>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "cosma_amd.h"
#include <cosma/memory_pool.hpp>
#include <cosma/mpi_types.hpp>
#include <cosma/strategy.hpp>

namespace cosma {

// The reference's communicator (src/cosma/communicator.hpp); a single rank
// never creates one.
class communicator;

template <typename Scalar>
class cosma_context {
  public:
    // Streams and tile sizes from COSMA_GPU_STREAMS / COSMA_GPU_MAX_TILE_*,
    // memory limit from COSMA_CPU_MAX_MEMORY, flags from the other knobs.
    cosma_context();
    cosma_context(std::size_t cpu_mem_limit, int streams, int tile_m, int tile_n, int tile_k);
    ~cosma_context();
    cosma_context(const cosma_context &) = delete;
    cosma_context &operator=(const cosma_context &) = delete;

    // Records the communicator and strategy of a multiply
    // (src/cosma/context.cpp:80-125): a new strategy drops the pool's
    // "already pinned" state, as in the reference.  The communicator must
    // have one rank.
    void register_state(MPI_Comm comm, const Strategy strategy);

    // The host memory of the CosmaMatrix objects of this context.
    memory_pool<Scalar> &get_memory_pool();

    // The device handle every local multiply of this context runs on
    // (created on first use, on the current HIP device 0).
    cosma_amd_handle_t get_gpu_context();

    // The reference's communicator of the last multiply; nullptr here: a
    // single rank communicates with no one, and no communicator is built.
    cosma::communicator *get_cosma_comm();

    long long get_cpu_memory_limit();
    void turn_on_output();
    bool unified_memory();

    // Not in the reference interface: the streams and tile sizes the device
    // handle is (or will be) created with (COSMA_GPU_STREAMS,
    // COSMA_GPU_MAX_TILE_{M,N,K}, or the explicit constructor's arguments).
    int get_gpu_streams() const { return streams_; }
    int get_gpu_tile_m() const { return tile_m_; }
    int get_gpu_tile_n() const { return tile_n_; }
    int get_gpu_tile_k() const { return tile_k_; }

    bool adapt_to_scalapack_strategy = true;
    bool overlap_comm_and_comp = false;
    bool pin_host_buffers = true;

  private:
    long long cpu_memory_limit = std::numeric_limits<long long>::max();
    memory_pool<Scalar> memory_pool_;
    bool have_strategy_ = false;
    long long prev_m_ = 0, prev_n_ = 0, prev_k_ = 0;
    std::vector<int> prev_divisors_;
    std::string prev_dims_, prev_types_;
    int streams_ = 2, tile_m_ = 5000, tile_n_ = 5000, tile_k_ = 5000;
    cosma_amd_handle_t gpu_ctx_ = nullptr;
    bool output = false;
    bool use_unified_memory_ = false;
};

template <typename Scalar>
using global_context = cosma_context<Scalar> *;

template <typename Scalar>
using context = std::unique_ptr<cosma_context<Scalar>>;

template <typename Scalar>
context<Scalar> make_context();

template <typename Scalar>
context<Scalar> make_context(std::size_t cpu_mem_limit, int streams, int tile_m, int tile_n, int tile_k);

// Process-wide context (constructed on first use; thread-safe initialisation).
template <typename Scalar>
global_context<Scalar> get_context_instance();

}  // namespace cosma
