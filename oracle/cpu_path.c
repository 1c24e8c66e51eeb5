/*
 * cpu_path.c -- SPLA's CPU path, restated for the CPU baseline and as a
 * second oracle. At one rank, spla_p?gemm_ssb / spla_?gemm on a host context
 * end in exactly one host-BLAS call (src/gemm/gemm_host.cpp:55-72 ->
 * src/util/blas_interface.cpp:147-192).
 *   if m == 0 or n == 0: return
 *   lda, ldb, ldc clamped to >= 1
 *   cblas_?gemm(CblasColMajor, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
 * The BLAS library is loaded at run time (dlopen): MKL's libmkl_rt if present,
 * else an OpenBLAS build; the caller names the library and its symbol prefix.
 */
#include <dlfcn.h>
#include <stddef.h>
#include <stdio.h>

enum { CBLAS_COL_MAJOR = 102, CBLAS_NO_TRANS = 111, CBLAS_TRANS = 112, CBLAS_CONJ_TRANS = 113 };

typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double*, int, const double*,
                         int, double, double*, int);
typedef void (*zgemm_fn)(int, int, int, int, int, int, const void*, const void*, int, const void*,
                         int, const void*, void*, int);
typedef void (*set_threads_fn)(int);
typedef int (*get_threads_fn)(void);

static dgemm_fn g_dgemm = NULL;
static zgemm_fn g_zgemm = NULL;
static set_threads_fn g_set_threads = NULL;
static get_threads_fn g_get_threads = NULL;

/* Loads the BLAS library; prefix is "" (MKL) or "scipy_" (OpenBLAS wheel). */
int cpu_path_init(const char* libPath, const char* prefix, const char* setThreadsSym,
                  const char* getThreadsSym) {
  void* h = dlopen(libPath, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "cpu_path: %s\n", dlerror());
    return 1;
  }
  char name[128];
  snprintf(name, sizeof(name), "%scblas_dgemm", prefix);
  g_dgemm = (dgemm_fn)dlsym(h, name);
  snprintf(name, sizeof(name), "%scblas_zgemm", prefix);
  g_zgemm = (zgemm_fn)dlsym(h, name);
  g_set_threads = setThreadsSym ? (set_threads_fn)dlsym(h, setThreadsSym) : NULL;
  g_get_threads = getThreadsSym ? (get_threads_fn)dlsym(h, getThreadsSym) : NULL;
  return (g_dgemm && g_zgemm) ? 0 : 2;
}

void cpu_path_set_threads(int n) {
  if (g_set_threads) g_set_threads(n);
}

int cpu_path_get_threads(void) { return g_get_threads ? g_get_threads() : -1; }

static int to_cblas(int op) {
  return op == 0 ? CBLAS_NO_TRANS : (op == 1 ? CBLAS_TRANS : CBLAS_CONJ_TRANS);
}

static int at_least_one(int v) { return v < 1 ? 1 : v; }

/* blas::gemm for double (blas_interface.cpp:124-145 via gemm_host.cpp:55-72). */
void cpu_path_dgemm(int opA, int opB, int m, int n, int k, double alpha, const double* A, int lda,
                    const double* B, int ldb, double beta, double* C, int ldc) {
  if (m == 0 || n == 0) return;
  g_dgemm(CBLAS_COL_MAJOR, to_cblas(opA), to_cblas(opB), m, n, k, alpha, A, at_least_one(lda), B,
          at_least_one(ldb), beta, C, at_least_one(ldc));
}

/* blas::gemm for complex double (blas_interface.cpp:170-192). */
void cpu_path_zgemm(int opA, int opB, int m, int n, int k, const double* alpha, const double* A,
                    int lda, const double* B, int ldb, const double* beta, double* C, int ldc) {
  if (m == 0 || n == 0) return;
  g_zgemm(CBLAS_COL_MAJOR, to_cblas(opA), to_cblas(opB), m, n, k, alpha, A, at_least_one(lda), B,
          at_least_one(ldb), beta, C, at_least_one(ldc));
}
