// Internal state of spla::MatrixDistribution.
//
// Validation follows SPLA's MatrixDistributionInternal
// (src/spla/matrix_distribution_internal.cpp:43-150): order must be one of
// 'R','r','C','c'; grid and block sizes >= 1; a mapping must name distinct
// ranks in [0, comm size). This engine runs one rank: the communicator is kept
// as given (no reordered communicator is created) and products on a
// communicator with more than one rank raise MPIError.
#pragma once

#include <mpi.h>

#include <vector>

#include "spla/types.h"

namespace spla {

class MatrixDistributionInternal {
public:
  static MatrixDistributionInternal create_blacs_block_cyclic(MPI_Comm comm, char order,
                                                             int procGridRows, int procGridCols,
                                                             int rowBlockSize, int colBlockSize);
  static MatrixDistributionInternal create_blacs_block_cyclic_from_mapping(
      MPI_Comm comm, const int* mapping, int procGridRows, int procGridCols, int rowBlockSize,
      int colBlockSize);
  static MatrixDistributionInternal create_mirror(MPI_Comm comm);

  int proc_grid_rows() const { return procGridRows_; }
  int proc_grid_cols() const { return procGridCols_; }
  int row_block_size() const { return rowBlockSize_; }
  int col_block_size() const { return colBlockSize_; }
  void set_row_block_size(int v) { rowBlockSize_ = v; }
  void set_col_block_size(int v) { colBlockSize_ = v; }
  SplaDistributionType type() const { return type_; }
  MPI_Comm comm() const { return comm_; }
  int comm_size() const { return commSize_; }
  // Column-major grid position -> rank.
  const std::vector<int>& mapping() const { return mapping_; }

private:
  MatrixDistributionInternal() = default;

  SplaDistributionType type_ = SPLA_DIST_BLACS_BLOCK_CYCLIC;
  int procGridRows_ = 1, procGridCols_ = 1;
  int rowBlockSize_ = 256, colBlockSize_ = 256;
  MPI_Comm comm_;
  int commSize_ = 1;
  std::vector<int> mapping_;
};

}  // namespace spla
