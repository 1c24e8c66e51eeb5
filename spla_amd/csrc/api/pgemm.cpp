// pgemm_ssb / pgemm_ssbtr / pgemm_sbs, C++ and C (SPLA src/spla/pgemm_ssb.cpp:48-111,
// pgemm_ssbtr.cpp:46-195, pgemm_sbs.cpp:47-194).
//
// On a one-rank communicator SPLA turns every distributed product into one
// local GEMM on the sub-matrix (pgemm_ssb_gpu.cpp:259-262,
// pgemm_sbs_gpu.cpp:246-249); this engine does the same and raises MPIError
// for communicators with more than one rank (the ring transport is not part
// of this library). pgemm_ssbtr skips tiles of C that lie wholly outside the
// requested triangle, which the API allows ("any other part may or may not
// be computed").
#include <complex>
#include <cstdint>

#include "api/internal_access.hpp"
#include "runtime/engine.hpp"
#include "runtime/timer.hpp"
#include "spla/pgemm_sbs.h"
#include "spla/pgemm_sbs.hpp"
#include "spla/pgemm_ssb.h"
#include "spla/pgemm_ssb.hpp"
#include "spla/pgemm_ssbtr.h"
#include "spla/pgemm_ssbtr.hpp"

namespace spla {
namespace {

// SPLA pgemm_ssb_gpu (src/pgemm_ssb/pgemm_ssb_gpu.cpp:242-262).
template <typename T>
void ssbtr_impl(int m, int n, int kLocal, SplaOperation opA, T alpha, const T* A, int lda,
                const T* B, int ldb, T beta, T* C, int ldc, int cRowOffset, int cColOffset,
                SplaFillMode cFillMode, MatrixDistribution& distC, Context& ctx) {
  if (m == 0 || n == 0) return;
  spla_amd::timer::LibraryScope scope("pgemm_ssb_gpu");
  if (opA != SPLA_OP_TRANSPOSE && opA != SPLA_OP_CONJ_TRANSPOSE) throw InvalidParameterError();
  if (m < 0 || n < 0 || cRowOffset < 0 || cColOffset < 0) throw InvalidParameterError();
  MatrixDistributionInternal& desc = InternalAccess::dist(distC);
  if (desc.comm_size() != 1) throw MPIError();
  spla_amd::FillSpec fill;
  fill.mode = (cFillMode == SPLA_FILL_MODE_UPPER || cFillMode == SPLA_FILL_MODE_LOWER)
                  ? cFillMode
                  : SPLA_FILL_MODE_FULL;
  fill.rowOff = cRowOffset;
  fill.colOff = cColOffset;
  spla_amd::gemm_gpu<T>(opA, SPLA_OP_NONE, m, n, kLocal, alpha, A, lda, B, ldb, beta,
                        C + cRowOffset + static_cast<int64_t>(cColOffset) * ldc, ldc,
                        InternalAccess::ctx(ctx), fill);
}

// SPLA pgemm_sbs_gpu (src/pgemm_sbs/pgemm_sbs_gpu.cpp:234-249).
template <typename T>
void sbs_impl(int mLocal, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb,
              int bRowOffset, int bColOffset, MatrixDistribution& distB, T beta, T* C, int ldc,
              Context& ctx) {
  if (n == 0 || k == 0) return;
  spla_amd::timer::LibraryScope scope("pgemm_sbs_gpu");
  if (n < 0 || k < 0 || bRowOffset < 0 || bColOffset < 0) throw InvalidParameterError();
  MatrixDistributionInternal& desc = InternalAccess::dist(distB);
  if (desc.comm_size() != 1 && desc.type() != SPLA_DIST_MIRROR) throw MPIError();
  spla_amd::gemm_gpu<T>(SPLA_OP_NONE, SPLA_OP_NONE, mLocal, n, k, alpha, A, lda,
                        B + bRowOffset + static_cast<int64_t>(bColOffset) * ldb, ldb, beta, C,
                        ldc, InternalAccess::ctx(ctx));
}

}  // namespace

#define SPLA_PGEMM_OVERLOADS(T)                                                                   \
  void pgemm_ssb(int m, int n, int kLocal, SplaOperation opA, T alpha, const T* A, int lda,        \
                 const T* B, int ldb, T beta, T* C, int ldc, int cRowOffset, int cColOffset,       \
                 MatrixDistribution& distC, Context& ctx) {                                        \
    ssbtr_impl<T>(m, n, kLocal, opA, alpha, A, lda, B, ldb, beta, C, ldc, cRowOffset, cColOffset,  \
                  SPLA_FILL_MODE_FULL, distC, ctx);                                                \
  }                                                                                                \
  void pgemm_ssbtr(int m, int n, int kLocal, SplaOperation opA, T alpha, const T* A, int lda,      \
                   const T* B, int ldb, T beta, T* C, int ldc, int cRowOffset, int cColOffset,     \
                   SplaFillMode cFillMode, MatrixDistribution& distC, Context& ctx) {              \
    ssbtr_impl<T>(m, n, kLocal, opA, alpha, A, lda, B, ldb, beta, C, ldc, cRowOffset, cColOffset,  \
                  cFillMode, distC, ctx);                                                          \
  }                                                                                                \
  void pgemm_sbs(int mLocal, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb,      \
                 int bRowOffset, int bColOffset, MatrixDistribution& descB, T beta, T* C, int ldc, \
                 Context& ctx) {                                                                   \
    sbs_impl<T>(mLocal, n, k, alpha, A, lda, B, ldb, bRowOffset, bColOffset, descB, beta, C, ldc,  \
                ctx);                                                                              \
  }

SPLA_PGEMM_OVERLOADS(float)
SPLA_PGEMM_OVERLOADS(double)
SPLA_PGEMM_OVERLOADS(std::complex<float>)
SPLA_PGEMM_OVERLOADS(std::complex<double>)
#undef SPLA_PGEMM_OVERLOADS

}  // namespace spla

namespace {
spla::Context& as_ctx(SplaContext c) { return *reinterpret_cast<spla::Context*>(c); }
spla::MatrixDistribution& as_dist(SplaMatrixDistribution d) {
  return *reinterpret_cast<spla::MatrixDistribution*>(d);
}
template <typename T>
T value_of(const void* p) {
  return *static_cast<const T*>(p);
}
}  // namespace

extern "C" {

// Real precisions take scalars by value, complex ones by pointer.
#define SPLA_C_SSB(NAME, T, ST, PT, CONV)                                                     \
  SPLA_EXPORT SplaError NAME(int m, int n, int kLocal, SplaOperation opA, ST alpha,           \
                             const PT* A, int lda, const PT* B, int ldb, ST beta, PT* C,       \
                             int ldc, int cRowOffset, int cColOffset,                         \
                             SplaMatrixDistribution distC, SplaContext ctx) {                 \
    if (!ctx || !distC) return SPLA_INVALID_HANDLE_ERROR;                                     \
    SPLA_C_GUARD(spla::pgemm_ssb(m, n, kLocal, opA, CONV(alpha), reinterpret_cast<const T*>(A), lda, \
                                 reinterpret_cast<const T*>(B), ldb, CONV(beta), reinterpret_cast<T*>(C), \
                                 ldc, cRowOffset, cColOffset, as_dist(distC), as_ctx(ctx)))   \
  }
#define SPLA_C_SSBTR(NAME, T, ST, PT, CONV)                                                     \
  SPLA_EXPORT SplaError NAME(int m, int n, int kLocal, SplaOperation opA, ST alpha,             \
                             const PT* A, int lda, const PT* B, int ldb, ST beta, PT* C,         \
                             int ldc, int cRowOffset, int cColOffset, SplaFillMode cFillMode,    \
                             SplaMatrixDistribution distC, SplaContext ctx) {                   \
    if (!ctx || !distC) return SPLA_INVALID_HANDLE_ERROR;                                       \
    SPLA_C_GUARD(spla::pgemm_ssbtr(m, n, kLocal, opA, CONV(alpha), reinterpret_cast<const T*>(A),    \
                                   lda, reinterpret_cast<const T*>(B), ldb, CONV(beta),              \
                                   reinterpret_cast<T*>(C), ldc, cRowOffset, cColOffset, cFillMode,   \
                                   as_dist(distC), as_ctx(ctx)))                                \
  }
#define SPLA_C_SBS(NAME, T, ST, PT, CONV)                                                       \
  SPLA_EXPORT SplaError NAME(int mLocal, int n, int k, ST alpha, const PT* A, int lda,          \
                             const PT* B, int ldb, int bRowOffset, int bColOffset,               \
                             SplaMatrixDistribution distB, ST beta, PT* C, int ldc,              \
                             SplaContext ctx) {                                                  \
    if (!ctx || !distB) return SPLA_INVALID_HANDLE_ERROR;                                       \
    SPLA_C_GUARD(spla::pgemm_sbs(mLocal, n, k, CONV(alpha), reinterpret_cast<const T*>(A), lda,      \
                                 reinterpret_cast<const T*>(B), ldb, bRowOffset, bColOffset,          \
                                 as_dist(distB), CONV(beta), reinterpret_cast<T*>(C), ldc,            \
                                 as_ctx(ctx)))                                                   \
  }

#define SPLA_ID(x) (x)
#define SPLA_CF(x) value_of<std::complex<float>>(x)
#define SPLA_CD(x) value_of<std::complex<double>>(x)

SPLA_C_SSB(spla_psgemm_ssb, float, float, float, SPLA_ID)
SPLA_C_SSB(spla_pdgemm_ssb, double, double, double, SPLA_ID)
SPLA_C_SSB(spla_pcgemm_ssb, std::complex<float>, const void*, void, SPLA_CF)
SPLA_C_SSB(spla_pzgemm_ssb, std::complex<double>, const void*, void, SPLA_CD)
SPLA_C_SSBTR(spla_psgemm_ssbtr, float, float, float, SPLA_ID)
SPLA_C_SSBTR(spla_pdgemm_ssbtr, double, double, double, SPLA_ID)
SPLA_C_SSBTR(spla_pcgemm_ssbtr, std::complex<float>, const void*, void, SPLA_CF)
SPLA_C_SSBTR(spla_pzgemm_ssbtr, std::complex<double>, const void*, void, SPLA_CD)
SPLA_C_SBS(spla_psgemm_sbs, float, float, float, SPLA_ID)
SPLA_C_SBS(spla_pdgemm_sbs, double, double, double, SPLA_ID)
SPLA_C_SBS(spla_pcgemm_sbs, std::complex<float>, const void*, void, SPLA_CF)
SPLA_C_SBS(spla_pzgemm_sbs, std::complex<double>, const void*, void, SPLA_CD)

}  // extern "C"
