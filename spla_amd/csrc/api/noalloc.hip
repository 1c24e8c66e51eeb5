// spla_amd_zgemm_noalloc: spla_zgemm for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). The FP64 kernels' split-K partial sums go to `scratch`,
// a file-scope device array that every GPU the library runs on holds until the
// process ends; nothing else of the device path needs device memory.
//
// Every call on one GPU shares that GPU's array and leaves its last contents
// there: calls that overlap on one GPU (two host threads, or two contexts on
// different streams) race. Nothing in it has to start at zero: the GEMM kernel
// writes every partial sum that the reduction kernel of the same call reads.
#include <hip/hip_runtime.h>

#include <complex>
#include <cstddef>
#include <mutex>

#include "api/internal_access.hpp"
#include "runtime/engine.hpp"
#include "runtime/hip_check.hpp"
#include "spla/context.hpp"
#include "spla_amd/spla_amd.h"

namespace {

constexpr std::size_t kScratchBytes = std::size_t(64) << 20;
constexpr int kMaxDevices = 64;

// Device-side scratch; in host code this name is only a handle for
// hipGetSymbolAddress, never an address to pass to a kernel or a copy.
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

// Per-device address of `scratch` (each GPU holds its own copy of the array).
void* scratch_address(int device) {
  static std::mutex lock;
  static void* address[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) throw spla::GPUInvalidValueError();
  std::lock_guard<std::mutex> guard(lock);
  if (!address[device]) {
    int prev = 0;
    spla_amd::hip_check(hipGetDevice(&prev));
    if (prev != device) spla_amd::hip_check(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    spla_amd::hip_check(st);
    address[device] = p;
  }
  return address[device];
}

}  // namespace

extern "C" {

SPLA_EXPORT SplaError spla_amd_zgemm_noalloc(SplaOperation opA, SplaOperation opB, int m, int n,
                                             int k, const void* alpha, const void* A, int lda,
                                             const void* B, int ldb, const void* beta, void* C,
                                             int ldc, SplaContext ctx) {
  if (!ctx) return SPLA_INVALID_HANDLE_ERROR;
  using T = std::complex<double>;
  SPLA_C_GUARD({
    spla::ContextInternal& c = spla::InternalAccess::ctx(*reinterpret_cast<spla::Context*>(ctx));
    if (m != 0 && n != 0) {
      void* ws = scratch_address(c.device());
      spla_amd::gemm_gpu_fixed_workspace<T>(
          opA, opB, m, n, k, *static_cast<const T*>(alpha), static_cast<const T*>(A), lda,
          static_cast<const T*>(B), ldb, *static_cast<const T*>(beta), static_cast<T*>(C), ldc, c,
          ws, kScratchBytes);
    }
  })
}

}  // extern "C"
