// cosma_context on one MI355X (interface of src/cosma/context.hpp:19-66;
// constructors src/cosma/context.cpp:15-52, destructor :54-62, singleton
// :154-158).  The device handle is created on first use so that merely
// constructing a context (e.g. the singleton at start-up) touches no GPU.
#include <iostream>

#include <cosma/context.hpp>
#include <cosma/environment_variables.hpp>

#include "front.hpp"

namespace cosma {

template <typename Scalar>
cosma_context<Scalar>::cosma_context() {
    cpu_memory_limit = get_cpu_max_memory<Scalar>();
    streams_ = gpu_streams();
    tile_m_ = gpu_max_tile_m();
    tile_n_ = gpu_max_tile_n();
    tile_k_ = gpu_max_tile_k();
    adapt_to_scalapack_strategy = get_adapt_strategy();
    overlap_comm_and_comp = get_overlap_comm_and_comp();
    pin_host_buffers = get_memory_pinning();
    use_unified_memory_ = get_unified_memory();
}

template <typename Scalar>
cosma_context<Scalar>::cosma_context(std::size_t cpu_mem_limit, int streams, int tile_m, int tile_n, int tile_k)
    : cosma_context() {
    cpu_memory_limit = (long long)cpu_mem_limit;
    streams_ = streams;
    tile_m_ = tile_m;
    tile_n_ = tile_n;
    tile_k_ = tile_k;
}

template <typename Scalar>
cosma_context<Scalar>::~cosma_context() {
    if (gpu_ctx_) cosma_amd_handle_destroy(gpu_ctx_);
}

template <typename Scalar>
void cosma_context<Scalar>::register_state(MPI_Comm comm, const Strategy strategy) {
    const int P = detail::comm_size(comm);
    if (P != 1)
        detail::fail("register_state: the communicator has " + std::to_string(P) + " ranks; this build runs one");
    const bool same = have_strategy_ && strategy.m == prev_m_ && strategy.n == prev_n_ && strategy.k == prev_k_ &&
                      strategy.divisors == prev_divisors_ && strategy.split_dimension == prev_dims_ &&
                      strategy.step_type == prev_types_;
    if (!same) {
        memory_pool_.unpin_all();
        have_strategy_ = true;
        prev_m_ = strategy.m;
        prev_n_ = strategy.n;
        prev_k_ = strategy.k;
        prev_divisors_ = strategy.divisors;
        prev_dims_ = strategy.split_dimension;
        prev_types_ = strategy.step_type;
    }
}

template <typename Scalar>
memory_pool<Scalar> &cosma_context<Scalar>::get_memory_pool() {
    return memory_pool_;
}

template <typename Scalar>
cosma::communicator *cosma_context<Scalar>::get_cosma_comm() {
    return nullptr;
}

template <typename Scalar>
cosma_amd_handle_t cosma_context<Scalar>::get_gpu_context() {
    if (!gpu_ctx_) {
        detail::check(cosma_amd_handle_create(&gpu_ctx_, 0, streams_, tile_m_, tile_n_, tile_k_),
                      "cosma_context");
        if (output)
            std::cout << "cosma_amd: device handle with " << streams_ << " streams, tiles " << tile_m_ << " x "
                      << tile_n_ << " x " << tile_k_ << std::endl;
    }
    return gpu_ctx_;
}

template <typename Scalar>
long long cosma_context<Scalar>::get_cpu_memory_limit() {
    return cpu_memory_limit;
}

template <typename Scalar>
void cosma_context<Scalar>::turn_on_output() {
    output = true;
    memory_pool_.turn_on_output();
}

template <typename Scalar>
bool cosma_context<Scalar>::unified_memory() {
    return use_unified_memory_;
}

template <typename Scalar>
context<Scalar> make_context() {
    return std::make_unique<cosma_context<Scalar>>();
}

template <typename Scalar>
context<Scalar> make_context(std::size_t cpu_mem_limit, int streams, int tile_m, int tile_n, int tile_k) {
    return std::make_unique<cosma_context<Scalar>>(cpu_mem_limit, streams, tile_m, tile_n, tile_k);
}

template <typename Scalar>
global_context<Scalar> get_context_instance() {
    static cosma_context<Scalar> instance;
    return &instance;
}

template class cosma_context<double>;
template context<double> make_context<double>();
template context<double> make_context<double>(std::size_t, int, int, int, int);
template global_context<double> get_context_instance<double>();

}  // namespace cosma
