/*
 * Stripe-block-stripe product: C <- alpha * A * B + beta * C where A
 * (mLocal x k) and C (mLocal x n) are row stripes and B (k x n) is
 * distributed by distB; (bRowOffset, bColOffset) locate the B sub-matrix.
 *
 * C entry points of SPLA 1.6.1 include/spla/pgemm_sbs.h:85-115 (implemented
 * in src/spla/pgemm_sbs.cpp:124-194).
 */
#ifndef SPLA_PGEMM_SBS_H
#define SPLA_PGEMM_SBS_H

#include "spla/config.h"
#include "spla/context.h"
#include "spla/errors.h"
#include "spla/matrix_distribution.h"

#ifdef __cplusplus
extern "C" {
#endif

SPLA_EXPORT SplaError spla_psgemm_sbs(int mLocal, int n, int k, float alpha, const float* A,
                                      int lda, const float* B, int ldb, int bRowOffset,
                                      int bColOffset, SplaMatrixDistribution distB, float beta,
                                      float* C, int ldc, SplaContext ctx);

SPLA_EXPORT SplaError spla_pdgemm_sbs(int mLocal, int n, int k, double alpha, const double* A,
                                      int lda, const double* B, int ldb, int bRowOffset,
                                      int bColOffset, SplaMatrixDistribution distB, double beta,
                                      double* C, int ldc, SplaContext ctx);

SPLA_EXPORT SplaError spla_pcgemm_sbs(int mLocal, int n, int k, const void* alpha, const void* A,
                                      int lda, const void* B, int ldb, int bRowOffset,
                                      int bColOffset, SplaMatrixDistribution distB,
                                      const void* beta, void* C, int ldc, SplaContext ctx);

SPLA_EXPORT SplaError spla_pzgemm_sbs(int mLocal, int n, int k, const void* alpha, const void* A,
                                      int lda, const void* B, int ldb, int bRowOffset,
                                      int bColOffset, SplaMatrixDistribution distB,
                                      const void* beta, void* C, int ldc, SplaContext ctx);

#ifdef __cplusplus
}
#endif

#endif
