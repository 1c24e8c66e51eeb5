// spla::gemm overloads and spla_{s,d,c,z}gemm (SPLA src/spla/gemm.cpp:41-172).
// Both processing units are served by the GPU engine.
#include <complex>

#include "api/internal_access.hpp"
#include "runtime/engine.hpp"
#include "spla/gemm.h"
#include "spla/gemm.hpp"

namespace spla {

#define SPLA_GEMM_OVERLOAD(T)                                                                   \
  void gemm(SplaOperation opA, SplaOperation opB, int m, int n, int k, T alpha, const T* A,      \
            int lda, const T* B, int ldb, T beta, T* C, int ldc, Context& ctx) {                 \
    spla_amd::gemm_gpu<T>(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,                \
                          InternalAccess::ctx(ctx));                                             \
  }

SPLA_GEMM_OVERLOAD(float)
SPLA_GEMM_OVERLOAD(double)
SPLA_GEMM_OVERLOAD(std::complex<float>)
SPLA_GEMM_OVERLOAD(std::complex<double>)
#undef SPLA_GEMM_OVERLOAD

}  // namespace spla

namespace {
spla::Context& as_ctx(SplaContext ctx) { return *reinterpret_cast<spla::Context*>(ctx); }
}  // namespace

extern "C" {

SPLA_EXPORT SplaError spla_sgemm(SplaOperation opA, SplaOperation opB, int m, int n, int k,
                                 float alpha, const float* A, int lda, const float* B, int ldb,
                                 float beta, float* C, int ldc, SplaContext ctx) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(spla::gemm(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, as_ctx(ctx)))
}

SPLA_EXPORT SplaError spla_dgemm(SplaOperation opA, SplaOperation opB, int m, int n, int k,
                                 double alpha, const double* A, int lda, const double* B, int ldb,
                                 double beta, double* C, int ldc, SplaContext ctx) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  SPLA_C_GUARD(spla::gemm(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, as_ctx(ctx)))
}

SPLA_EXPORT SplaError spla_cgemm(SplaOperation opA, SplaOperation opB, int m, int n, int k,
                                 const void* alpha, const void* A, int lda, const void* B, int ldb,
                                 const void* beta, void* C, int ldc, SplaContext ctx) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  using T = std::complex<float>;
  SPLA_C_GUARD(spla::gemm(opA, opB, m, n, k, *static_cast<const T*>(alpha),
                          static_cast<const T*>(A), lda, static_cast<const T*>(B), ldb,
                          *static_cast<const T*>(beta), static_cast<T*>(C), ldc, as_ctx(ctx)))
}

SPLA_EXPORT SplaError spla_zgemm(SplaOperation opA, SplaOperation opB, int m, int n, int k,
                                 const void* alpha, const void* A, int lda, const void* B, int ldb,
                                 const void* beta, void* C, int ldc, SplaContext ctx) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  using T = std::complex<double>;
  SPLA_C_GUARD(spla::gemm(opA, opB, m, n, k, *static_cast<const T*>(alpha),
                          static_cast<const T*>(A), lda, static_cast<const T*>(B), ldb,
                          *static_cast<const T*>(beta), static_cast<T*>(C), ldc, as_ctx(ctx)))
}

}  // extern "C"
