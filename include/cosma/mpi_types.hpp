// MPI_Comm for the COSMA interface.  The interface takes a communicator, as
// the reference's does; this build runs one rank and never calls MPI itself
// (it asks MPI_Comm_size through the process's MPI only when one is loaded).
// With <mpi.h> available its MPI_Comm is used; otherwise the MPICH ABI's
// integer handle stands in.
#pragma once
#if __has_include(<mpi.h>)
#include <mpi.h>
#else
typedef int MPI_Comm;
#endif
