#include "runtime/matrix_distribution_internal.hpp"

#include <set>

#include "runtime/mpi_shim.hpp"
#include "spla/exceptions.hpp"

namespace spla {

MatrixDistributionInternal MatrixDistributionInternal::create_blacs_block_cyclic(
    MPI_Comm comm, char order, int procGridRows, int procGridCols, int rowBlockSize,
    int colBlockSize) {
  if (order != 'R' && order != 'r' && order != 'C' && order != 'c') throw InvalidParameterError();
  if (procGridRows < 1 || procGridCols < 1 || rowBlockSize < 1 || colBlockSize < 1)
    throw InvalidParameterError();
  // Column-major grid storage; 'R' numbers ranks along grid rows.
  std::vector<int> mapping(static_cast<std::size_t>(procGridRows) * procGridCols);
  for (int r = 0; r < procGridRows; ++r)
    for (int c = 0; c < procGridCols; ++c) {
      const int rank = (order == 'R' || order == 'r') ? r * procGridCols + c : c * procGridRows + r;
      mapping[static_cast<std::size_t>(c) * procGridRows + r] = rank;
    }
  return create_blacs_block_cyclic_from_mapping(comm, mapping.data(), procGridRows, procGridCols,
                                                rowBlockSize, colBlockSize);
}

MatrixDistributionInternal MatrixDistributionInternal::create_blacs_block_cyclic_from_mapping(
    MPI_Comm comm, const int* mapping, int procGridRows, int procGridCols, int rowBlockSize,
    int colBlockSize) {
  if (procGridRows < 1 || procGridCols < 1 || rowBlockSize < 1 || colBlockSize < 1)
    throw InvalidParameterError();
  if (!mapping) throw InvalidParameterError();
  const int size = spla_amd::comm_size(comm);
  const long long gridSize = static_cast<long long>(procGridRows) * procGridCols;
  std::set<int> seen;
  for (long long i = 0; i < gridSize; ++i) {
    if (mapping[i] < 0 || mapping[i] >= size) throw InvalidParameterError();
    if (!seen.insert(mapping[i]).second) throw InvalidParameterError();
  }
  MatrixDistributionInternal d;
  d.type_ = SPLA_DIST_BLACS_BLOCK_CYCLIC;
  d.procGridRows_ = procGridRows;
  d.procGridCols_ = procGridCols;
  d.rowBlockSize_ = rowBlockSize;
  d.colBlockSize_ = colBlockSize;
  d.comm_ = comm;
  d.commSize_ = size;
  d.mapping_.assign(mapping, mapping + gridSize);
  return d;
}

MatrixDistributionInternal MatrixDistributionInternal::create_mirror(MPI_Comm comm) {
  // SPLA: grid rows = comm size, one column, 256 x 256 blocks
  // (matrix_distribution_internal.cpp:152-168).
  MatrixDistributionInternal d;
  d.type_ = SPLA_DIST_MIRROR;
  d.comm_ = comm;
  d.commSize_ = spla_amd::comm_size(comm);
  d.procGridRows_ = d.commSize_;
  d.procGridCols_ = 1;
  d.rowBlockSize_ = 256;
  d.colBlockSize_ = 256;
  return d;
}

}  // namespace spla
