/*
 * C++ MatrixDistribution (SPLA 1.6.1 include/spla/matrix_distribution.hpp:47-222).
 * Copies share state; the object holds one shared pointer, as in SPLA.
 */
#ifndef SPLA_MATRIX_DISTRIBUTION_HPP
#define SPLA_MATRIX_DISTRIBUTION_HPP

#include <mpi.h>

#include <complex>
#include <memory>

#include "spla/config.h"
#include "spla/types.h"

namespace spla {

class MatrixDistributionInternal;
struct InternalAccess;

class SPLA_EXPORT MatrixDistribution {
public:
  /* Block cyclic layout; order 'R'/'r' numbers grid ranks row-major, 'C'/'c' column-major. */
  static MatrixDistribution create_blacs_block_cyclic(MPI_Comm comm, char order, int procGridRows,
                                                      int procGridCols, int rowBlockSize,
                                                      int colBlockSize);

  /* Block cyclic layout with an explicit column-major grid -> rank mapping. */
  static MatrixDistribution create_blacs_block_cyclic_from_mapping(
      MPI_Comm comm, const int* mapping, int procGridRows, int procGridCols, int rowBlockSize,
      int colBlockSize);

  /* Every rank holds the full matrix. */
  static MatrixDistribution create_mirror(MPI_Comm comm);

  MatrixDistribution(MatrixDistribution&&) = default;
  MatrixDistribution(const MatrixDistribution&) = default;
  MatrixDistribution& operator=(MatrixDistribution&&) = default;
  MatrixDistribution& operator=(const MatrixDistribution&) = default;

  int proc_grid_rows() const;
  int proc_grid_cols() const;
  int row_block_size() const;
  int col_block_size() const;
  SplaDistributionType type() const;
  MPI_Comm comm();
  void set_row_block_size(int rowBlockSize);
  void set_col_block_size(int colBlockSize);

private:
  explicit MatrixDistribution(std::shared_ptr<MatrixDistributionInternal> descInternal);
  friend struct InternalAccess;
  std::shared_ptr<MatrixDistributionInternal> descInternal_;
};

}  // namespace spla

#endif
