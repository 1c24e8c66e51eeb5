// spla::MatrixDistribution and the spla_mat_dis_* C entry points
// (SPLA src/spla/matrix_distribution.cpp:36-267).
#include "api/internal_access.hpp"
#include "spla/matrix_distribution.h"
#include "spla/matrix_distribution.hpp"

namespace spla {

MatrixDistribution MatrixDistribution::create_blacs_block_cyclic(MPI_Comm comm, char order,
                                                                 int procGridRows,
                                                                 int procGridCols,
                                                                 int rowBlockSize,
                                                                 int colBlockSize) {
  return InternalAccess::make_dist(MatrixDistributionInternal::create_blacs_block_cyclic(
      comm, order, procGridRows, procGridCols, rowBlockSize, colBlockSize));
}

MatrixDistribution MatrixDistribution::create_blacs_block_cyclic_from_mapping(
    MPI_Comm comm, const int* mapping, int procGridRows, int procGridCols, int rowBlockSize,
    int colBlockSize) {
  return InternalAccess::make_dist(
      MatrixDistributionInternal::create_blacs_block_cyclic_from_mapping(
          comm, mapping, procGridRows, procGridCols, rowBlockSize, colBlockSize));
}

MatrixDistribution MatrixDistribution::create_mirror(MPI_Comm comm) {
  return InternalAccess::make_dist(MatrixDistributionInternal::create_mirror(comm));
}

MatrixDistribution::MatrixDistribution(std::shared_ptr<MatrixDistributionInternal> d)
    : descInternal_(std::move(d)) {}

int MatrixDistribution::proc_grid_rows() const { return descInternal_->proc_grid_rows(); }
int MatrixDistribution::proc_grid_cols() const { return descInternal_->proc_grid_cols(); }
int MatrixDistribution::row_block_size() const { return descInternal_->row_block_size(); }
int MatrixDistribution::col_block_size() const { return descInternal_->col_block_size(); }
SplaDistributionType MatrixDistribution::type() const { return descInternal_->type(); }
MPI_Comm MatrixDistribution::comm() { return descInternal_->comm(); }
void MatrixDistribution::set_row_block_size(int v) { descInternal_->set_row_block_size(v); }
void MatrixDistribution::set_col_block_size(int v) { descInternal_->set_col_block_size(v); }

}  // namespace spla

namespace {
spla::MatrixDistribution& as_dist(SplaMatrixDistribution d) {
  return *reinterpret_cast<spla::MatrixDistribution*>(d);
}
}  // namespace

extern "C" {

SPLA_EXPORT SplaError spla_mat_dis_create_block_cyclic(SplaMatrixDistribution* matDis,
                                                       MPI_Comm comm, char order, int procGridRows,
                                                       int procGridCols, int rowBlockSize,
                                                       int colBlockSize) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(*matDis = reinterpret_cast<void*>(
                   new spla::MatrixDistribution(spla::MatrixDistribution::create_blacs_block_cyclic(
                       comm, order, procGridRows, procGridCols, rowBlockSize, colBlockSize))))
}

SPLA_EXPORT SplaError spla_mat_dis_create_block_cyclic_fortran(SplaMatrixDistribution* matDis,
                                                               int commFortran, char order,
                                                               int procGridRows, int procGridCols,
                                                               int rowBlockSize, int colBlockSize) {
  return spla_mat_dis_create_block_cyclic(matDis, MPI_Comm_f2c(commFortran), order, procGridRows,
                                          procGridCols, rowBlockSize, colBlockSize);
}

SPLA_EXPORT SplaError spla_mat_dis_create_blacs_block_cyclic_from_mapping(
    SplaMatrixDistribution* matDis, MPI_Comm comm, const int* mapping, int procGridRows,
    int procGridCols, int rowBlockSize, int colBlockSize) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(*matDis = reinterpret_cast<void*>(new spla::MatrixDistribution(
                   spla::MatrixDistribution::create_blacs_block_cyclic_from_mapping(
                       comm, mapping, procGridRows, procGridCols, rowBlockSize, colBlockSize))))
}

SPLA_EXPORT SplaError spla_mat_dis_create_blacs_block_cyclic_from_mapping_fortran(
    SplaMatrixDistribution* matDis, int commFortran, const int* mapping, int procGridRows,
    int procGridCols, int rowBlockSize, int colBlockSize) {
  return spla_mat_dis_create_blacs_block_cyclic_from_mapping(
      matDis, MPI_Comm_f2c(commFortran), mapping, procGridRows, procGridCols, rowBlockSize,
      colBlockSize);
}

SPLA_EXPORT SplaError spla_mat_dis_destroy(SplaMatrixDistribution* matDis) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  try {
    delete reinterpret_cast<spla::MatrixDistribution*>(*matDis);
  } catch (const spla::GenericError& e) {
    return e.error_code();
  } catch (...) {
    return SPLA_UNKNOWN_ERROR;
  }
  *matDis = nullptr;
  return SPLA_SUCCESS;
}

SPLA_EXPORT SplaError spla_mat_dis_create_mirror(SplaMatrixDistribution* matDis, MPI_Comm comm) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(*matDis = reinterpret_cast<void*>(
                   new spla::MatrixDistribution(spla::MatrixDistribution::create_mirror(comm))))
}

SPLA_EXPORT SplaError spla_mat_dis_create_mirror_fortran(SplaMatrixDistribution* matDis,
                                                         int commFortran) {
  return spla_mat_dis_create_mirror(matDis, MPI_Comm_f2c(commFortran));
}

#define SPLA_DIS_GETTER(NAME, TYPE, EXPR)                             \
  SPLA_EXPORT SplaError NAME(SplaMatrixDistribution matDis, TYPE* out) { \
    if (!matDis) return SPLA_INVALID_HANDLE_ERROR;                     \
    SPLA_C_GUARD(*out = as_dist(matDis).EXPR)                          \
  }

SPLA_DIS_GETTER(spla_mat_dis_proc_grid_rows, int, proc_grid_rows())
SPLA_DIS_GETTER(spla_mat_dis_proc_grid_cols, int, proc_grid_cols())
SPLA_DIS_GETTER(spla_mat_dis_row_block_size, int, row_block_size())
SPLA_DIS_GETTER(spla_mat_dis_col_block_size, int, col_block_size())
SPLA_DIS_GETTER(spla_mat_dis_type, SplaDistributionType, type())
SPLA_DIS_GETTER(spla_mat_dis_comm, MPI_Comm, comm())
#undef SPLA_DIS_GETTER

SPLA_EXPORT SplaError spla_mat_dis_comm_fortran(SplaMatrixDistribution matDis, int* commFortran) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(*commFortran = MPI_Comm_c2f(as_dist(matDis).comm()))
}

SPLA_EXPORT SplaError spla_mat_dis_set_row_block_size(SplaMatrixDistribution matDis,
                                                      int rowBlockSize) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(as_dist(matDis).set_row_block_size(rowBlockSize))
}

SPLA_EXPORT SplaError spla_mat_dis_set_col_block_size(SplaMatrixDistribution matDis,
                                                      int colBlockSize) {
  if (!matDis) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(as_dist(matDis).set_col_block_size(colBlockSize))
}

}  // extern "C"
