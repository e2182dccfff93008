// Internal helpers of the C++ front-end (the COSMA interface on top of the
// cosma_amd C ABI).
#pragma once
#include <stdexcept>
#include <string>

#include "cosma_amd.h"
#include <cosma/mpi_types.hpp>

namespace cosma {
namespace detail {

// Throws std::runtime_error("<where>: <library error text>") if status != 0.
inline void check(int status, const char *where) {
    if (status != COSMA_AMD_OK) throw std::runtime_error(std::string(where) + ": " + cosma_amd_last_error());
}

[[noreturn]] inline void fail(const std::string &what) { throw std::runtime_error(what); }

// Size of the communicator when an MPICH-ABI MPI is loaded and initialised in
// the process, else 1.
int comm_size(MPI_Comm comm);

// Looks a symbol up in the process (its first definition in the global
// lookup order); nullptr if absent.
void *find_symbol(const char *name);

}  // namespace detail
}  // namespace cosma
