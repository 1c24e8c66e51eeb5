/*
 * Enumerations of the public API. Ordinals are part of the C ABI and match
 * SPLA 1.6.1 include/spla/types.h:33-83.
 */
#ifndef SPLA_TYPES_H
#define SPLA_TYPES_H

#include "spla/config.h"

/* How a matrix is spread over the ranks of a communicator. */
enum SplaDistributionType {
  SPLA_DIST_BLACS_BLOCK_CYCLIC, /* 0: BLACS-style block cyclic layout */
  SPLA_DIST_MIRROR              /* 1: every rank holds the full matrix */
};

/* Where a context executes its work. */
enum SplaProcessingUnit {
  SPLA_PU_HOST, /* 0: host CPU */
  SPLA_PU_GPU   /* 1: GPU */
};

/* Operation applied to an operand when it is read. */
enum SplaOperation {
  SPLA_OP_NONE,          /* 0: as stored */
  SPLA_OP_TRANSPOSE,     /* 1: transposed */
  SPLA_OP_CONJ_TRANSPOSE /* 2: conjugate transposed */
};

/* Part of C that must be computed by pgemm_ssbtr. */
enum SplaFillMode {
  SPLA_FILL_MODE_FULL,  /* 0: whole matrix */
  SPLA_FILL_MODE_UPPER, /* 1: upper triangle (including diagonal) */
  SPLA_FILL_MODE_LOWER  /* 2: lower triangle (including diagonal) */
};

#ifndef __cplusplus
typedef enum SplaDistributionType SplaDistributionType;
typedef enum SplaProcessingUnit SplaProcessingUnit;
typedef enum SplaOperation SplaOperation;
typedef enum SplaFillMode SplaFillMode;
#endif

#endif
