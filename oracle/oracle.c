/*
 * oracle.c -- CPU restatement of SPLA's local-GEMM path, used as the test
 * oracle and nothing else (only tests/, the smoke check and the bench's CPU
 * baseline load it; the library never does).
 *
 * SPLA's CPU path at one rank is one host-BLAS call:
 *   spla_p?gemm_ssb -> pgemm_ssb_host -> gemm_host -> blas::gemm -> cblas_?gemm
 *   (src/pgemm_ssb/pgemm_ssb_host.cpp:168-171, src/gemm/gemm_host.cpp:55-72,
 *    src/util/blas_interface.cpp:147-192).
 * Its semantics are those of the reference BLAS ?GEMM (column-major
 * C <- alpha op(A) op(B) + beta C; m == 0 or n == 0 returns; k == 0 or
 * alpha == 0 gives C <- beta C without reading A or B; beta == 0 does not
 * read C). This file restates them twice:
 *   oracle_gemm_{d,z}    plain double accumulation in index order;
 *   oracle_gemm_{d,z}_dd double-double accumulation (error-free products by
 *                        fma, Knuth TwoSum), used as the FP64-accurate
 *                        reference for error measurement;
 * plus SPLA's parameter check (src/util/check_gemm_param.hpp:36-50).
 *
 * Parity pins (tests/test_oracle.py): exact-rational golden vectors
 * (tests/golden/), host BLAS on the same inputs, integer-valued inputs for
 * which every result is exact, the BLAS special cases above, and planted
 * mistakes (wrong conjugation, transposed operand, dropped term) that the
 * golden vectors must reject.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

/* Operation codes: SplaOperation ordinals (include/spla/types.h:55-67). */
enum { OP_N = 0, OP_T = 1, OP_C = 2 };

typedef struct {
  double re, im;
} zdouble;

/* ---- element access of op(X) ---------------------------------------- */

static inline double d_op(const double* X, int64_t ld, int op, int64_t r, int64_t c) {
  return op == OP_N ? X[r + c * ld] : X[c + r * ld];
}

static inline zdouble z_op(const zdouble* X, int64_t ld, int op, int64_t r, int64_t c) {
  zdouble v = op == OP_N ? X[r + c * ld] : X[c + r * ld];
  if (op == OP_C) v.im = -v.im;
  return v;
}

/* ---- plain restatement ---------------------------------------------- */

void oracle_gemm_d(int opA, int opB, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                   double* C, int64_t ldc) {
  if (m == 0 || n == 0) return;
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < m; ++i) {
      double acc = 0.0;
      if (alpha != 0.0)
        for (int64_t l = 0; l < k; ++l) acc += d_op(A, lda, opA, i, l) * d_op(B, ldb, opB, l, j);
      double out = alpha == 0.0 ? 0.0 : alpha * acc;
      if (beta != 0.0) out += beta * C[i + j * ldc];
      C[i + j * ldc] = out;
    }
  }
}

void oracle_gemm_z(int opA, int opB, int64_t m, int64_t n, int64_t k, const double* alpha,
                   const double* A_, int64_t lda, const double* B_, int64_t ldb,
                   const double* beta, double* C_, int64_t ldc) {
  const zdouble* A = (const zdouble*)A_;
  const zdouble* B = (const zdouble*)B_;
  zdouble* C = (zdouble*)C_;
  const int alphaZero = alpha[0] == 0.0 && alpha[1] == 0.0;
  const int betaZero = beta[0] == 0.0 && beta[1] == 0.0;
  if (m == 0 || n == 0) return;
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < m; ++i) {
      double re = 0.0, im = 0.0;
      if (!alphaZero)
        for (int64_t l = 0; l < k; ++l) {
          const zdouble a = z_op(A, lda, opA, i, l);
          const zdouble b = z_op(B, ldb, opB, l, j);
          re += a.re * b.re - a.im * b.im;
          im += a.re * b.im + a.im * b.re;
        }
      zdouble out = {0.0, 0.0};
      if (!alphaZero) {
        out.re = alpha[0] * re - alpha[1] * im;
        out.im = alpha[0] * im + alpha[1] * re;
      }
      if (!betaZero) {
        const zdouble c = C[i + j * ldc];
        out.re += beta[0] * c.re - beta[1] * c.im;
        out.im += beta[0] * c.im + beta[1] * c.re;
      }
      C[i + j * ldc] = out;
    }
  }
}

/* ---- double-double restatement --------------------------------------- */

typedef struct {
  double hi, lo;
} dd;

static inline dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  dd r = {s, e};
  return r;
}

static inline dd dd_add(dd x, dd y) {
  dd s = two_sum(x.hi, y.hi);
  s.lo += x.lo + y.lo;
  return two_sum(s.hi, s.lo);
}

/* exact product a*b as an unevaluated sum (fma-based TwoProd) */
static inline dd two_prod(double a, double b) {
  const double p = a * b;
  dd r = {p, fma(a, b, -p)};
  return r;
}

static inline dd dd_mul_d(dd x, double b) {
  dd p = two_prod(x.hi, b);
  p.lo += x.lo * b;
  return two_sum(p.hi, p.lo);
}

static inline dd dd_neg(dd x) {
  dd r = {-x.hi, -x.lo};
  return r;
}

void oracle_gemm_d_dd(int opA, int opB, int64_t m, int64_t n, int64_t k, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                      double* C, int64_t ldc) {
  if (m == 0 || n == 0) return;
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < m; ++i) {
      dd acc = {0.0, 0.0};
      if (alpha != 0.0)
        for (int64_t l = 0; l < k; ++l)
          acc = dd_add(acc, two_prod(d_op(A, lda, opA, i, l), d_op(B, ldb, opB, l, j)));
      dd out = {0.0, 0.0};
      if (alpha != 0.0) out = dd_mul_d(acc, alpha);
      if (beta != 0.0) out = dd_add(out, two_prod(beta, C[i + j * ldc]));
      C[i + j * ldc] = out.hi + out.lo;
    }
  }
}

/* Accurate complex entries of alpha op(A) op(B) + beta C at selected (row, col)
 * positions; out receives (re, im) pairs. With rows == NULL every entry of C is
 * computed in place. */
static void z_dd_entry(int opA, int opB, int64_t k, const double* alpha, const zdouble* A,
                       int64_t lda, const zdouble* B, int64_t ldb, const double* beta,
                       const zdouble* cIn, int64_t i, int64_t j, double* outRe, double* outIm) {
  const int alphaZero = alpha[0] == 0.0 && alpha[1] == 0.0;
  const int betaZero = beta[0] == 0.0 && beta[1] == 0.0;
  dd re = {0.0, 0.0}, im = {0.0, 0.0};
  if (!alphaZero)
    for (int64_t l = 0; l < k; ++l) {
      const zdouble a = z_op(A, lda, opA, i, l);
      const zdouble b = z_op(B, ldb, opB, l, j);
      re = dd_add(re, two_prod(a.re, b.re));
      re = dd_add(re, dd_neg(two_prod(a.im, b.im)));
      im = dd_add(im, two_prod(a.re, b.im));
      im = dd_add(im, two_prod(a.im, b.re));
    }
  dd oRe = {0.0, 0.0}, oIm = {0.0, 0.0};
  if (!alphaZero) {
    oRe = dd_add(dd_mul_d(re, alpha[0]), dd_neg(dd_mul_d(im, alpha[1])));
    oIm = dd_add(dd_mul_d(im, alpha[0]), dd_mul_d(re, alpha[1]));
  }
  if (!betaZero) {
    oRe = dd_add(oRe, dd_add(two_prod(beta[0], cIn->re), dd_neg(two_prod(beta[1], cIn->im))));
    oIm = dd_add(oIm, dd_add(two_prod(beta[0], cIn->im), two_prod(beta[1], cIn->re)));
  }
  *outRe = oRe.hi + oRe.lo;
  *outIm = oIm.hi + oIm.lo;
}

void oracle_gemm_z_dd(int opA, int opB, int64_t m, int64_t n, int64_t k, const double* alpha,
                      const double* A_, int64_t lda, const double* B_, int64_t ldb,
                      const double* beta, double* C_, int64_t ldc) {
  const zdouble* A = (const zdouble*)A_;
  const zdouble* B = (const zdouble*)B_;
  zdouble* C = (zdouble*)C_;
  if (m == 0 || n == 0) return;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double re, im;
      z_dd_entry(opA, opB, k, alpha, A, lda, B, ldb, beta, &C[i + j * ldc], i, j, &re, &im);
      C[i + j * ldc].re = re;
      C[i + j * ldc].im = im;
    }
}

/* Sampled entries: out[2*s], out[2*s+1] = entry (rows[s], cols[s]) of the
 * result; cIn holds the initial C (may be NULL when beta == 0). */
void oracle_gemm_z_dd_entries(int opA, int opB, int64_t k, const double* alpha, const double* A_,
                              int64_t lda, const double* B_, int64_t ldb, const double* beta,
                              const double* cIn_, int64_t ldc, const int64_t* rows,
                              const int64_t* cols, int64_t count, double* out) {
  const zdouble zero = {0.0, 0.0};
  for (int64_t s = 0; s < count; ++s) {
    const zdouble* c = cIn_ ? (const zdouble*)cIn_ + rows[s] + cols[s] * ldc : &zero;
    z_dd_entry(opA, opB, k, alpha, (const zdouble*)A_, lda, (const zdouble*)B_, ldb, beta, c,
               rows[s], cols[s], &out[2 * s], &out[2 * s + 1]);
  }
}

void oracle_gemm_d_dd_entries(int opA, int opB, int64_t k, double alpha, const double* A,
                              int64_t lda, const double* B, int64_t ldb, double beta,
                              const double* cIn, int64_t ldc, const int64_t* rows,
                              const int64_t* cols, int64_t count, double* out) {
  for (int64_t s = 0; s < count; ++s) {
    const int64_t i = rows[s], j = cols[s];
    dd acc = {0.0, 0.0};
    if (alpha != 0.0)
      for (int64_t l = 0; l < k; ++l)
        acc = dd_add(acc, two_prod(d_op(A, lda, opA, i, l), d_op(B, ldb, opB, l, j)));
    dd o = {0.0, 0.0};
    if (alpha != 0.0) o = dd_mul_d(acc, alpha);
    if (beta != 0.0 && cIn) o = dd_add(o, two_prod(beta, cIn[i + j * ldc]));
    out[s] = o.hi + o.lo;
  }
}

/* ---- SPLA parameter check --------------------------------------------- */

/* check_gemm_param (src/util/check_gemm_param.hpp:36-50). Returns the
 * SplaError code: 0 success, 3 invalid parameter, 4 invalid pointer. */
int oracle_check_gemm_param(int opA, int opB, int64_t m, int64_t n, int64_t k, const void* A,
                            int64_t lda, const void* B, int64_t ldb, const void* C) {
  if (m < 0 || n < 0 || k < 0) return 3;
  if ((opA == OP_N && lda < m) || (opA != OP_N && lda < k)) return 3;
  if ((opB == OP_N && ldb < k) || (opB != OP_N && ldb < n)) return 3;
  if (k != 0 && m != 0 && !A) return 4;
  if (k != 0 && n != 0 && !B) return 4;
  if (n != 0 && m != 0 && !C) return 4;
  return 0;
}
