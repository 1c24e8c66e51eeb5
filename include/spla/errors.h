/*
 * Error codes returned by every C entry point. Ordinals are part of the C ABI
 * and match SPLA 1.6.1 include/spla/errors.h:33-106.
 */
#ifndef SPLA_ERRORS_H
#define SPLA_ERRORS_H

#include "spla/config.h"

enum SplaError {
  SPLA_SUCCESS,                          /*  0 */
  SPLA_UNKNOWN_ERROR,                    /*  1 */
  SPLA_INTERNAL_ERROR,                   /*  2 */
  SPLA_INVALID_PARAMETER_ERROR,          /*  3 */
  SPLA_INVALID_POINTER_ERROR,            /*  4 */
  SPLA_INVALID_HANDLE_ERROR,             /*  5 */
  SPLA_MPI_ERROR,                        /*  6 */
  SPLA_MPI_ALLOCATION_ERROR,             /*  7 */
  SPLA_MPI_THREAD_SUPPORT_ERROR,         /*  8 */
  SPLA_GPU_ERROR,                        /*  9 */
  SPLA_GPU_SUPPORT_ERROR,                /* 10 */
  SPLA_GPU_ALLOCATION_ERROR,             /* 11 */
  SPLA_GPU_LAUNCH_ERROR,                 /* 12 */
  SPLA_GPU_NO_DEVICE_ERROR,              /* 13 */
  SPLA_GPU_INVALID_VALUE_ERROR,          /* 14 */
  SPLA_GPU_INVALID_DEVICE_POINTER_ERROR, /* 15 */
  SPLA_GPU_BLAS_ERROR,                   /* 16 */
  SPLA_INVALID_ALLOCATOR_FUNCTION        /* 17 */
};

#ifndef __cplusplus
typedef enum SplaError SplaError;
#endif

#endif
