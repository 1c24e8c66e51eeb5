/*
 * Stripe-stripe-block product: C <- alpha * op(A) * B + beta * C, where A
 * (kLocal x m) and B (kLocal x n) are row stripes, op(A) is the transpose or
 * conjugate transpose of A (opA must be SPLA_OP_TRANSPOSE or
 * SPLA_OP_CONJ_TRANSPOSE) and C is distributed by distC.
 *
 * C entry points of SPLA 1.6.1 include/spla/pgemm_ssb.h:88-121 (implemented in
 * src/spla/pgemm_ssb.cpp:81-111, which forwards to the ssbtr variant with
 * SPLA_FILL_MODE_FULL).
 */
#ifndef SPLA_PGEMM_SSB_H
#define SPLA_PGEMM_SSB_H

#include "spla/config.h"
#include "spla/context.h"
#include "spla/errors.h"
#include "spla/matrix_distribution.h"

#ifdef __cplusplus
extern "C" {
#endif

SPLA_EXPORT SplaError spla_psgemm_ssb(int m, int n, int kLocal, SplaOperation opA, float alpha,
                                      const float* A, int lda, const float* B, int ldb, float beta,
                                      float* C, int ldc, int cRowOffset, int cColOffset,
                                      SplaMatrixDistribution distC, SplaContext ctx);

SPLA_EXPORT SplaError spla_pdgemm_ssb(int m, int n, int kLocal, SplaOperation opA, double alpha,
                                      const double* A, int lda, const double* B, int ldb,
                                      double beta, double* C, int ldc, int cRowOffset,
                                      int cColOffset, SplaMatrixDistribution distC,
                                      SplaContext ctx);

SPLA_EXPORT SplaError spla_pcgemm_ssb(int m, int n, int kLocal, SplaOperation opA,
                                      const void* alpha, const void* A, int lda, const void* B,
                                      int ldb, const void* beta, void* C, int ldc, int cRowOffset,
                                      int cColOffset, SplaMatrixDistribution distC,
                                      SplaContext ctx);

SPLA_EXPORT SplaError spla_pzgemm_ssb(int m, int n, int kLocal, SplaOperation opA,
                                      const void* alpha, const void* A, int lda, const void* B,
                                      int ldb, const void* beta, void* C, int ldc, int cRowOffset,
                                      int cColOffset, SplaMatrixDistribution distC,
                                      SplaContext ctx);

#ifdef __cplusplus
}
#endif

#endif
