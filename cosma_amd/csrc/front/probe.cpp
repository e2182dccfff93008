// Run-time lookups of MPI and other libraries the process may have loaded.
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <dlfcn.h>

#include "front.hpp"

namespace cosma {
namespace detail {

void *find_symbol(const char *name) { return dlsym(RTLD_DEFAULT, name); }

int comm_size(MPI_Comm comm) {
#if !defined(OPEN_MPI)
    // Open MPI's handles are pointers, not the integer handles this
    // interface was built with: nothing can be asked then (build with
    // Open MPI's mpi.h on the include path to ask it).
    if (find_symbol("ompi_mpi_comm_world")) return 1;
#endif
    using initialized_fn = int (*)(int *);
    using size_fn = int (*)(MPI_Comm, int *);
    auto initialized = reinterpret_cast<initialized_fn>(find_symbol("MPI_Initialized"));
    auto size = reinterpret_cast<size_fn>(find_symbol("MPI_Comm_size"));
    if (!initialized || !size) return 1;
    int flag = 0;
    if (initialized(&flag) != 0 || !flag) return 1;
    int n = 1;
    if (size(comm, &n) != 0) return 1;
    return n;
}

}  // namespace detail
}  // namespace cosma
