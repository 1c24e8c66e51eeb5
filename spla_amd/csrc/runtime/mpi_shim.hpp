// Communicator queries for the single-rank engine.
//
// The library does not link an MPI implementation. MPI functions are looked
// up in the process (dlsym with RTLD_DEFAULT) the first time they are needed:
// a program that calls MPI_Init has them, a program without MPI has not. The
// MPI_Comm handle type is the one of <mpi.h> the library is compiled against
// (MPICH ABI: a 32-bit integer).
//   * MPI present and initialised: MPI_Comm_size of the communicator.
//   * MPI absent or not initialised: there is only one process, so every
//     communicator has size 1.
#pragma once

#include <mpi.h>

namespace spla_amd {

// Number of ranks in comm (see above). Throws spla::MPIError on MPI failure.
int comm_size(MPI_Comm comm);

// Rank of the caller in comm (0 without MPI).
int comm_rank(MPI_Comm comm);

// True when an MPI library is present in the process and initialised.
bool mpi_active();

}  // namespace spla_amd
