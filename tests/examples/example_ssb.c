/*
 * Drop-in check of the C API: a C program written against SPLA's headers
 * computes C = A^H B with spla_pzgemm_ssb on MPI_COMM_SELF and compares a
 * naive host product. Usage: example_ssb [m n k]. Prints "max error E" and
 * returns 0 when E <= 1e-12 (relative to the largest |C| entry). <mpi.h>
 * provides the communicator type and MPI_COMM_SELF; one rank needs no MPI
 * runtime (the library queries MPI only when it is present and initialised).
 */
#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>

#include "spla/spla.h"

#define CHECK(call)                                                   \
  do {                                                                \
    SplaError e_ = (call);                                            \
    if (e_ != SPLA_SUCCESS) {                                         \
      fprintf(stderr, "%s failed with SplaError %d\n", #call, (int)e_); \
      return 2;                                                       \
    }                                                                 \
  } while (0)

static double value(long i, int which) { return sin(0.37 * (double)i + 1.3 * which) * 0.5; }

int main(int argc, char** argv) {
  int m = argc > 3 ? atoi(argv[1]) : 97;
  int n = argc > 3 ? atoi(argv[2]) : 61;
  int k = argc > 3 ? atoi(argv[3]) : 1500;
  double* A = malloc(sizeof(double) * 2 * (size_t)k * m);
  double* B = malloc(sizeof(double) * 2 * (size_t)k * n);
  double* C = malloc(sizeof(double) * 2 * (size_t)m * n);
  for (long i = 0; i < 2L * k * m; ++i) A[i] = value(i, 1);
  for (long i = 0; i < 2L * k * n; ++i) B[i] = value(i, 2);

  SplaContext ctx = NULL;
  CHECK(spla_ctx_create(&ctx, SPLA_PU_GPU));
  SplaMatrixDistribution dist = NULL;
  CHECK(spla_mat_dis_create_block_cyclic(&dist, MPI_COMM_SELF, 'R', 1, 1, 64, 64));
  const double alpha[2] = {1.0, 0.0}, beta[2] = {0.0, 0.0};
  CHECK(spla_pzgemm_ssb(m, n, k, SPLA_OP_CONJ_TRANSPOSE, alpha, A, k, B, k, beta, C, m, 0, 0,
                        dist, ctx));

  double maxErr = 0.0, maxVal = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double re = 0.0, im = 0.0;
      for (int l = 0; l < k; ++l) {
        const double ar = A[2 * (l + (long)i * k)], ai = -A[2 * (l + (long)i * k) + 1];
        const double br = B[2 * (l + (long)j * k)], bi = B[2 * (l + (long)j * k) + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      const double* c = &C[2 * (i + (long)j * m)];
      maxErr = fmax(maxErr, fmax(fabs(c[0] - re), fabs(c[1] - im)));
      maxVal = fmax(maxVal, fmax(fabs(re), fabs(im)));
    }
  const double rel = maxErr / maxVal;
  printf("max error %.3e\n", rel);
  CHECK(spla_mat_dis_destroy(&dist));
  CHECK(spla_ctx_destroy(&ctx));
  free(A);
  free(B);
  free(C);
  return rel <= 1e-12 ? 0 : 1;
}
