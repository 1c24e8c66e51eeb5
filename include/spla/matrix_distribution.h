/*
 * Matrix distributions (block cyclic over a process grid, or mirrored).
 *
 * C entry points of SPLA 1.6.1 include/spla/matrix_distribution.h:43-186,
 * implemented there in src/spla/matrix_distribution.cpp:88-267. The
 * *_fortran variants take a Fortran communicator handle
 * (matrix_distribution.cpp:103,128,163,257).
 */
#ifndef SPLA_MATRIX_DISTRIBUTION_H
#define SPLA_MATRIX_DISTRIBUTION_H

#include <mpi.h>

#include "spla/config.h"
#include "spla/errors.h"
#include "spla/types.h"

/* Opaque handle holding a heap-allocated spla::MatrixDistribution. */
typedef void* SplaMatrixDistribution;

#ifdef __cplusplus
extern "C" {
#endif

/* Block cyclic distribution; order 'R'/'r' numbers grid ranks row-major, 'C'/'c' column-major. */
SPLA_EXPORT SplaError spla_mat_dis_create_block_cyclic(SplaMatrixDistribution* matDis,
                                                       MPI_Comm comm, char order, int procGridRows,
                                                       int procGridCols, int rowBlockSize,
                                                       int colBlockSize);

/* Block cyclic distribution with an explicit column-major grid -> rank mapping. */
SPLA_EXPORT SplaError spla_mat_dis_create_blacs_block_cyclic_from_mapping(
    SplaMatrixDistribution* matDis, MPI_Comm comm, const int* mapping, int procGridRows,
    int procGridCols, int rowBlockSize, int colBlockSize);

/* Destroy a distribution and set *matDis to NULL; NULL -> SPLA_INVALID_HANDLE_ERROR. */
SPLA_EXPORT SplaError spla_mat_dis_destroy(SplaMatrixDistribution* matDis);

/* Mirror distribution: every rank stores the whole matrix. */
SPLA_EXPORT SplaError spla_mat_dis_create_mirror(SplaMatrixDistribution* matDis, MPI_Comm comm);

SPLA_EXPORT SplaError spla_mat_dis_proc_grid_rows(SplaMatrixDistribution matDis, int* procGridRows);
SPLA_EXPORT SplaError spla_mat_dis_proc_grid_cols(SplaMatrixDistribution matDis, int* procGridCols);
SPLA_EXPORT SplaError spla_mat_dis_row_block_size(SplaMatrixDistribution matDis, int* rowBlockSize);
SPLA_EXPORT SplaError spla_mat_dis_col_block_size(SplaMatrixDistribution matDis, int* colBlockSize);
SPLA_EXPORT SplaError spla_mat_dis_type(SplaMatrixDistribution matDis, SplaDistributionType* type);

/* Communicator used internally (rank order may differ from the one given at creation). */
SPLA_EXPORT SplaError spla_mat_dis_comm(SplaMatrixDistribution matDis, MPI_Comm* comm);

SPLA_EXPORT SplaError spla_mat_dis_set_row_block_size(SplaMatrixDistribution matDis,
                                                      int rowBlockSize);
SPLA_EXPORT SplaError spla_mat_dis_set_col_block_size(SplaMatrixDistribution matDis,
                                                      int colBlockSize);

/* Fortran-handle variants, used by SPLA's Fortran module (spla.f90). */
SPLA_EXPORT SplaError spla_mat_dis_create_block_cyclic_fortran(SplaMatrixDistribution* matDis,
                                                               int commFortran, char order,
                                                               int procGridRows, int procGridCols,
                                                               int rowBlockSize, int colBlockSize);
SPLA_EXPORT SplaError spla_mat_dis_create_blacs_block_cyclic_from_mapping_fortran(
    SplaMatrixDistribution* matDis, int commFortran, const int* mapping, int procGridRows,
    int procGridCols, int rowBlockSize, int colBlockSize);
SPLA_EXPORT SplaError spla_mat_dis_create_mirror_fortran(SplaMatrixDistribution* matDis,
                                                         int commFortran);
SPLA_EXPORT SplaError spla_mat_dis_comm_fortran(SplaMatrixDistribution matDis, int* commFortran);

#ifdef __cplusplus
}
#endif

#endif
