#include "runtime/mpi_shim.hpp"

#include <dlfcn.h>

#include "spla/exceptions.hpp"

namespace spla_amd {
namespace {

using InitializedFn = int (*)(int*);
using FinalizedFn = int (*)(int*);
using CommQueryFn = int (*)(MPI_Comm, int*);

template <typename Fn>
Fn lookup(const char* name) {
  return reinterpret_cast<Fn>(dlsym(RTLD_DEFAULT, name));
}

}  // namespace

bool mpi_active() {
  auto initialized = lookup<InitializedFn>("MPI_Initialized");
  auto finalized = lookup<FinalizedFn>("MPI_Finalized");
  if (!initialized) return false;
  int flag = 0;
  if (initialized(&flag) != MPI_SUCCESS || !flag) return false;
  if (finalized) {
    int done = 0;
    if (finalized(&done) == MPI_SUCCESS && done) return false;
  }
  return true;
}

int comm_size(MPI_Comm comm) {
  if (!mpi_active()) return 1;
  auto fn = lookup<CommQueryFn>("MPI_Comm_size");
  if (!fn) return 1;
  int size = 0;
  if (fn(comm, &size) != MPI_SUCCESS) throw spla::MPIError();
  return size;
}

int comm_rank(MPI_Comm comm) {
  if (!mpi_active()) return 0;
  auto fn = lookup<CommQueryFn>("MPI_Comm_rank");
  if (!fn) return 0;
  int rank = 0;
  if (fn(comm, &rank) != MPI_SUCCESS) throw spla::MPIError();
  return rank;
}

}  // namespace spla_amd
