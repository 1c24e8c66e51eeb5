#include "runtime/engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "kernels/launch.hpp"
#include "runtime/hip_check.hpp"
#include "runtime/split_gemm.hpp"
#include "runtime/timer.hpp"
#include "spla/exceptions.hpp"

namespace spla_amd {

void check_gemm_param(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                      const void* A, int64_t lda, const void* B, int64_t ldb, const void* C,
                      int64_t ldc) {
  (void)ldc;  // SPLA does not check ldc
  if (m < 0 || n < 0 || k < 0) throw spla::InvalidParameterError();
  if ((opA == SPLA_OP_NONE && lda < m) || (opA != SPLA_OP_NONE && lda < k))
    throw spla::InvalidParameterError();
  if ((opB == SPLA_OP_NONE && ldb < k) || (opB != SPLA_OP_NONE && ldb < n))
    throw spla::InvalidParameterError();
  if (k != 0 && m != 0 && !A) throw spla::InvalidPointerError();
  if (k != 0 && n != 0 && !B) throw spla::InvalidPointerError();
  if (n != 0 && m != 0 && !C) throw spla::InvalidPointerError();
}

namespace {

enum class MemKind { Device, DmaHost, PageableHost };

// SPLA translate_gpu_pointer (src/gpu_util/gpu_pointer_translation.hpp:41-81):
// only device memory is used in place. Pinned and managed memory can be the
// source of a DMA copy; unregistered memory is staged through pinned buffers.
MemKind classify(const void* p) {
  if (!p) return MemKind::PageableHost;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  const hipError_t st = hipPointerGetAttributes(&attr, p);
  if (st != hipSuccess) {
    (void)hipGetLastError();
    return MemKind::PageableHost;
  }
  switch (attr.type) {
    case hipMemoryTypeDevice:
      return MemKind::Device;
    case hipMemoryTypeHost:
    case hipMemoryTypeManaged:
      return MemKind::DmaHost;
    default:
      return MemKind::PageableHost;
  }
}

class DeviceGuard {
public:
  explicit DeviceGuard(int dev) {
    hip_check(hipGetDevice(&prev_));
    if (prev_ != dev) {
      hip_check(hipSetDevice(dev));
      changed_ = true;
    }
  }
  ~DeviceGuard() {
    if (changed_) (void)hipSetDevice(prev_);
  }
  DeviceGuard(const DeviceGuard&) = delete;
  DeviceGuard& operator=(const DeviceGuard&) = delete;

private:
  int prev_ = 0;
  bool changed_ = false;
};

class Event {
public:
  Event() { hip_check(hipEventCreateWithFlags(&e_, hipEventDisableTiming)); }
  ~Event() { (void)hipEventDestroy(e_); }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  hipEvent_t get() const { return e_; }

private:
  hipEvent_t e_ = nullptr;
};

template <typename T>
struct Scalar;
template <>
struct Scalar<double> {
  static constexpr bool kComplex = false;
  static constexpr bool kF32 = false;
  using Real = double;
  static double re(double a) { return a; }
  static double im(double) { return 0.0; }
};
template <>
struct Scalar<std::complex<double>> {
  static constexpr bool kComplex = true;
  static constexpr bool kF32 = false;
  using Real = double;
  static double re(std::complex<double> a) { return a.real(); }
  static double im(std::complex<double> a) { return a.imag(); }
};
template <>
struct Scalar<float> {
  static constexpr bool kComplex = false;
  static constexpr bool kF32 = true;
  using Real = float;
  static double re(float a) { return a; }
  static double im(float) { return 0.0; }
};
template <>
struct Scalar<std::complex<float>> {
  static constexpr bool kComplex = true;
  static constexpr bool kF32 = true;
  using Real = float;
  static double re(std::complex<float> a) { return a.real(); }
  static double im(std::complex<float> a) { return a.imag(); }
};

// One product on device-resident operands, enqueued on `stream`. Workspace
// blocks are appended to `keep` and must live until the stream is drained.
// Skinny kernel: products with at least this embedded reduction length, split so
// that every split holds at least kSkinnyMinKpPerSplit (16 steps per wave),
// kSkinnyWgPerCU workgroups per CU.
constexpr int64_t kSkinnyMinKp = 8192;
constexpr int64_t kSkinnyMinKpPerSplit = 1024;
constexpr int kSkinnyWgPerCU = 4;
static_assert(kSkinnyMinKp >= 2 * kSkinnyMinKpPerSplit, "the skinny kernel writes split partials");

// `fixed` (optional): caller-provided split-K workspace; nothing is allocated,
// and a plan whose partial sums exceed it raises GPUAllocationError.
struct FixedWorkspace {
  void* ptr;
  std::size_t bytes;
};

template <typename T>
void enqueue_device_product(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                            T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                            T* C, int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                            hipStream_t stream, std::vector<Block>& keep, Block* reuse = nullptr,
                            bool trySplit = true, bool wholeChip = true,
                            const FixedWorkspace* fixed = nullptr) {
  using S = Scalar<T>;
  constexpr bool cplx = S::kComplex;
  constexpr bool f32 = S::kF32;
  GemmDesc d;
  std::memset(&d, 0, sizeof(d));
  d.A = A;
  d.B = B;
  d.C = C;
  d.lda = lda;
  d.ldb = ldb;
  d.ldc = ldc;
  d.m = static_cast<int>(m);
  d.n = static_cast<int>(n);
  d.k = k;
  d.Mp = static_cast<int>(cplx ? 2 * m : m);
  d.Kp = cplx ? 2 * k : k;
  d.alphaRe = S::re(alpha);
  d.alphaIm = S::im(alpha);
  d.betaRe = S::re(beta);
  d.betaIm = S::im(beta);
  d.betaZero = (d.betaRe == 0.0 && d.betaIm == 0.0) ? 1 : 0;
  d.conjA = (cplx && opA == SPLA_OP_CONJ_TRANSPOSE) ? 1 : 0;
  d.conjB = (cplx && opB == SPLA_OP_CONJ_TRANSPOSE) ? 1 : 0;
  d.fillMode = fill.mode;
  d.fillRowOff = fill.rowOff;
  d.fillColOff = fill.colOff;
  d.splits = 1;

  ctx.last_split() = SplitReport();
  if (k == 0 || (d.alphaRe == 0.0 && d.alphaIm == 0.0)) {
    // BLAS semantics: C <- beta * C, A and B are not read.
    hip_check(launch_scale(d, cplx, stream, f32));
    return;
  }

  const int la = (opA == SPLA_OP_NONE) ? MNMajor : KMajor;
  const int lb = (opB == SPLA_OP_NONE) ? KMajor : MNMajor;
  GemmPlan plan = ctx.tuner().plan(cplx, la, lb, m, n, k, ctx.num_cus(), wholeChip);
  // Small m, n with a long reduction on a whole device: HBM-bound, streamed by
  // the skinny kernel (one workgroup per split, no LDS staging).
  // (kSkinnyMinKp / kSkinnyMinKpPerSplit >= 2: always several splits)
  const bool skinny = !f32 && wholeChip && skinny_supported(cplx, la, lb, m, n) &&
                      d.Kp >= kSkinnyMinKp;
  if (skinny) {
    const int64_t maxSplits = std::max<int64_t>(1, d.Kp / kSkinnyMinKpPerSplit);
    int64_t sp = std::min<int64_t>(static_cast<int64_t>(ctx.num_cus()) * kSkinnyWgPerCU, maxSplits);
    const int64_t kSplit = (d.Kp + sp - 1) / sp;
    plan.kSplit = (kSplit + 15) / 16 * 16;
    plan.splits = static_cast<int>((d.Kp + plan.kSplit - 1) / plan.kSplit);
    plan.cfg = TileM;  // tile grid of the reduction's fill predicate
    plan.tilesM = plan.tilesN = 1;
  }
  if constexpr (!f32) {
    if (trySplit && !fixed && ctx.gemm_mode() != GemmModeFP64 &&
        enqueue_split_product<T>(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx,
                                 fill, plan.estSeconds, stream, keep, reuse))
      return;
  }
  d.splits = plan.splits;
  d.kSplit = plan.kSplit;
  d.tilesM = plan.tilesM;
  d.tilesN = plan.tilesN;
  if (plan.splits > 1) {
    const std::size_t bytes =
        static_cast<std::size_t>(plan.splits) * static_cast<std::size_t>(n) * d.Mp *
        sizeof(typename S::Real);
    if (fixed) {
      if (bytes > fixed->bytes) throw spla::GPUAllocationError();
      d.work = fixed->ptr;
    } else if (reuse) {
      // Products enqueued on one stream may share one workspace.
      if (reuse->bytes() < bytes) {
        keep.emplace_back(std::move(*reuse));
        *reuse = Block(ctx.gpu_memory(), bytes);
      }
      d.work = reuse->get();
    } else {
      keep.emplace_back(ctx.gpu_memory(), bytes);
      d.work = keep.back().get();
    }
  }
  {
    PhaseScope ps(ctx.timer(), PhaseGemmF64, stream);
    if (f32)
      hip_check(launch_gemm_f32(d, cplx, la, lb, plan.cfg, stream));
    else if (skinny)
      hip_check(launch_skinny_gemm(d, cplx, stream));
    else
      hip_check(launch_gemm_f64(d, cplx, la, lb, plan.cfg, stream));
  }
  if (plan.splits > 1) {
    PhaseScope ps(ctx.timer(), PhaseSplitKReduce, stream);
    hip_check(launch_splitk_reduce(d, cplx, plan.cfg, stream, f32));
  }
}

// Copies `cols` columns of `rowBytes` bytes (source pitch srcPitch) into a
// packed destination with host threads.
void pack_columns(void* dst, const void* src, std::size_t rowBytes, std::size_t srcPitch,
                  int64_t cols) {
  const std::size_t total = rowBytes * static_cast<std::size_t>(cols);
  unsigned hw = std::thread::hardware_concurrency();
  int nThreads = static_cast<int>(std::min<std::size_t>(std::max(1u, std::min(hw, 16u)),
                                                         total / (8u << 20) + 1));
  auto work = [&](int t) {
    const int64_t c0 = cols * t / nThreads;
    const int64_t c1 = cols * (t + 1) / nThreads;
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    if (rowBytes == srcPitch) {
      std::memcpy(d + c0 * rowBytes, s + c0 * srcPitch, (c1 - c0) * rowBytes);
    } else {
      for (int64_t c = c0; c < c1; ++c) std::memcpy(d + c * rowBytes, s + c * srcPitch, rowBytes);
    }
  };
  if (nThreads <= 1 || cols < 2) {
    nThreads = 1;
    work(0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nThreads - 1);
  for (int t = 1; t < nThreads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
}

// Sub-matrix of an operand covering reduction indices [k0, k0 + kc).
struct OperandChunk {
  const char* src;      // first element
  std::size_t rowBytes; // contiguous bytes per column of the sub-matrix
  std::size_t pitch;    // source leading dimension in bytes
  int64_t cols;         // number of columns
  int64_t packedLd;     // leading dimension (elements) once packed
};

template <typename T>
OperandChunk chunk_of(const T* X, int64_t ldx, bool kMajor, int64_t outerDim, int64_t k0,
                      int64_t kc) {
  OperandChunk c;
  if (kMajor) {  // stored (k x outer): rows [k0, k0+kc) of every column
    c.src = reinterpret_cast<const char*>(X + k0);
    c.rowBytes = static_cast<std::size_t>(kc) * sizeof(T);
    c.cols = outerDim;
    c.packedLd = kc;
  } else {  // stored (outer x k): columns [k0, k0+kc)
    c.src = reinterpret_cast<const char*>(X + k0 * ldx);
    c.rowBytes = static_cast<std::size_t>(outerDim) * sizeof(T);
    c.cols = kc;
    c.packedLd = outerDim;
  }
  c.pitch = static_cast<std::size_t>(ldx) * sizeof(T);
  return c;
}

template <typename T>
void gemm_gpu_f64(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k, T alpha,
                  const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc,
                  spla::ContextInternal& ctx, const FillSpec& fill) {
  const MemKind kindA = (k > 0) ? classify(A) : MemKind::Device;
  const MemKind kindB = (k > 0) ? classify(B) : MemKind::Device;
  const MemKind kindC = classify(C);
  const hipStream_t stream = ctx.compute_stream();
  const hipStream_t cstream = ctx.copy_stream();
  std::vector<Block> keep;

  if (kindA == MemKind::Device && kindB == MemKind::Device && kindC == MemKind::Device) {
    enqueue_device_product(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx, fill,
                           stream, keep);
    hip_check(hipStreamSynchronize(stream));
    return;
  }

  const bool betaZero = (beta == T(0));

  // C on the device for the whole call.
  T* dC = C;
  int64_t dldc = ldc;
  if (kindC != MemKind::Device) {
    keep.emplace_back(ctx.gpu_memory(), static_cast<std::size_t>(m) * n * sizeof(T));
    dC = keep.back().get<T>();
    dldc = m;
    // With a fill mode, tiles outside the triangle are not written but the whole
    // of dC is copied back: it must hold the caller's C there even when beta == 0.
    if (!betaZero || fill.mode != SPLA_FILL_MODE_FULL)
      hip_check(hipMemcpy2DAsync(dC, m * sizeof(T), C, ldc * sizeof(T), m * sizeof(T), n,
                                 hipMemcpyHostToDevice, stream));
  }

  const bool hostA = kindA != MemKind::Device;
  const bool hostB = kindB != MemKind::Device;
  if (!hostA && !hostB) {
    enqueue_device_product(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, dC, dldc, ctx, fill,
                           stream, keep);
  } else {
    // k-chunked streaming of the host-resident operands.
    const bool kMajorA = opA != SPLA_OP_NONE;
    const bool kMajorB = opB == SPLA_OP_NONE;
    const std::size_t bytesPerK = ((hostA ? m : 0) + (hostB ? n : 0)) * sizeof(T);
    const std::size_t targetBytes = std::max<std::size_t>(1, ctx.stream_chunk_bytes());
    int64_t kc = std::max<int64_t>(64, static_cast<int64_t>(targetBytes / bytesPerK));
    kc = std::min<int64_t>(kc, k);
    if (kc < k) kc = std::max<int64_t>(64, kc / 64 * 64);
    const int64_t numChunks = (k + kc - 1) / kc;
    constexpr int kSlots = 2;

    Block devA[kSlots], devB[kSlots], pinA[kSlots], pinB[kSlots];
    for (int s = 0; s < kSlots && s < numChunks; ++s) {
      if (hostA) {
        devA[s] = Block(ctx.gpu_memory(), kc * m * sizeof(T));
        if (kindA == MemKind::PageableHost) pinA[s] = Block(ctx.pinned_memory(), kc * m * sizeof(T));
      }
      if (hostB) {
        devB[s] = Block(ctx.gpu_memory(), kc * n * sizeof(T));
        if (kindB == MemKind::PageableHost) pinB[s] = Block(ctx.pinned_memory(), kc * n * sizeof(T));
      }
    }
    Event copied[kSlots], computed[kSlots];
    Block workspace;

    auto stage = [&](const T* X, int64_t ldx, bool kMajor, int64_t outer, MemKind kind,
                     Block& dev, Block& pin, int64_t k0, int64_t kcur, int slot) {
      const OperandChunk ch = chunk_of(X, ldx, kMajor, outer, k0, kcur);
      hip_check(hipStreamWaitEvent(cstream, computed[slot].get(), 0));
      if (kind == MemKind::DmaHost) {
        hip_check(hipMemcpy2DAsync(dev.get(), ch.rowBytes, ch.src, ch.pitch, ch.rowBytes, ch.cols,
                                   hipMemcpyHostToDevice, cstream));
      } else {
        pack_columns(pin.get(), ch.src, ch.rowBytes, ch.pitch, ch.cols);
        hip_check(hipMemcpyAsync(dev.get(), pin.get(), ch.rowBytes * ch.cols,
                                 hipMemcpyHostToDevice, cstream));
      }
      return ch.packedLd;
    };

    for (int64_t c = 0; c < numChunks; ++c) {
      const int slot = static_cast<int>(c % kSlots);
      const int64_t k0 = c * kc;
      const int64_t kcur = std::min(kc, k - k0);
      // The pinned slot is reused once the copy of chunk c - kSlots has finished.
      if (c >= kSlots) hip_check(hipEventSynchronize(copied[slot].get()));
      const T* cA = A;
      int64_t cLda = lda;
      const T* cB = B;
      int64_t cLdb = ldb;
      if (hostA) {
        cLda = stage(A, lda, kMajorA, m, kindA, devA[slot], pinA[slot], k0, kcur, slot);
        cA = devA[slot].get<T>();
      } else {
        cA = kMajorA ? A + k0 : A + k0 * lda;
      }
      if (hostB) {
        cLdb = stage(B, ldb, kMajorB, n, kindB, devB[slot], pinB[slot], k0, kcur, slot);
        cB = devB[slot].get<T>();
      } else {
        cB = kMajorB ? B + k0 : B + k0 * ldb;
      }
      hip_check(hipEventRecord(copied[slot].get(), cstream));
      hip_check(hipStreamWaitEvent(stream, copied[slot].get(), 0));
      // SPLA multiply_gpu applies beta to the first chunk only (multiply_gpu.cpp:141).
      const T betaChunk = (c == 0) ? beta : T(1);
      enqueue_device_product(opA, opB, m, n, kcur, alpha, cA, cLda, cB, cLdb, betaChunk, dC, dldc,
                             ctx, fill, stream, keep, &workspace);
      hip_check(hipEventRecord(computed[slot].get(), stream));
    }
  }

  if (kindC != MemKind::Device)
    hip_check(hipMemcpy2DAsync(C, ldc * sizeof(T), dC, m * sizeof(T), m * sizeof(T), n,
                               hipMemcpyDeviceToHost, stream));
  hip_check(hipStreamSynchronize(stream));
  hip_check(hipStreamSynchronize(cstream));
}

// Single precision on the FP32 matrix cores (FP32 accumulation, as the
// vendor s/c gemm SPLA calls). Host-resident operands are staged to the device
// once (packed), C is copied in (beta != 0) and out when it is host memory.
template <typename TF>
void gemm_gpu_f32(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                  TF alpha, const TF* A, int64_t lda, const TF* B, int64_t ldb, TF beta, TF* C,
                  int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill) {
  const hipStream_t stream = ctx.compute_stream();
  const int64_t rowsA = (opA == SPLA_OP_NONE) ? m : k, colsA = (opA == SPLA_OP_NONE) ? k : m;
  const int64_t rowsB = (opB == SPLA_OP_NONE) ? k : n, colsB = (opB == SPLA_OP_NONE) ? n : k;
  std::vector<Block> keep;

  auto on_device = [&](const TF* X, int64_t ldx, int64_t rows, int64_t cols,
                       int64_t& ldOut) -> const TF* {
    ldOut = ldx;
    if (rows * cols == 0 || classify(X) == MemKind::Device) return X;
    keep.emplace_back(ctx.gpu_memory(), static_cast<std::size_t>(rows) * cols * sizeof(TF));
    hip_check(hipMemcpy2DAsync(keep.back().get(), rows * sizeof(TF), X, ldx * sizeof(TF),
                               rows * sizeof(TF), cols, hipMemcpyHostToDevice, stream));
    ldOut = rows;
    return keep.back().get<TF>();
  };

  int64_t ldA = lda, ldB = ldb;
  const TF* dA = k > 0 ? on_device(A, lda, rowsA, colsA, ldA) : A;
  const TF* dB = k > 0 ? on_device(B, ldb, rowsB, colsB, ldB) : B;
  const bool hostC = classify(C) != MemKind::Device;
  TF* dC = C;
  int64_t ldC = ldc;
  if (hostC) {
    keep.emplace_back(ctx.gpu_memory(), static_cast<std::size_t>(m) * n * sizeof(TF));
    dC = keep.back().get<TF>();
    ldC = m;
    if (!(beta == TF(0)) || fill.mode != SPLA_FILL_MODE_FULL)  // see gemm_gpu_f64
      hip_check(hipMemcpy2DAsync(dC, m * sizeof(TF), C, ldc * sizeof(TF), m * sizeof(TF), n,
                                 hipMemcpyHostToDevice, stream));
  }
  enqueue_device_product<TF>(opA, opB, m, n, k, alpha, dA, ldA, dB, ldB, beta, dC, ldC, ctx, fill,
                             stream, keep);
  if (hostC)
    hip_check(hipMemcpy2DAsync(C, ldc * sizeof(TF), dC, m * sizeof(TF), m * sizeof(TF), n,
                               hipMemcpyDeviceToHost, stream));
  hip_check(hipStreamSynchronize(stream));
}

}  // namespace

template <typename T>
void gemm_gpu(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k, T alpha,
              const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc,
              spla::ContextInternal& ctx, FillSpec fill) {
  if (m == 0 || n == 0) return;
  timer::LibraryScope scope("gemm_gpu");
  check_gemm_param(opA, opB, m, n, k, A, lda, B, ldb, C, ldc);
  if (opA != SPLA_OP_NONE && opA != SPLA_OP_TRANSPOSE && opA != SPLA_OP_CONJ_TRANSPOSE)
    throw spla::InvalidParameterError();
  if (opB != SPLA_OP_NONE && opB != SPLA_OP_TRANSPOSE && opB != SPLA_OP_CONJ_TRANSPOSE)
    throw spla::InvalidParameterError();
  // The kernels index rows and columns of C with 32-bit integers.
  if (m > (int64_t(1) << 30) || n > (int64_t(1) << 30)) throw spla::InvalidParameterError();

  DeviceGuard guard(ctx.device());
  // SPLA synchronizes with the legacy default stream on entry (gemm_gpu.cpp:72).
  hip_check(hipStreamSynchronize(nullptr));
  ctx.timer().reset();
  struct Resolve {
    spla::ContextInternal& c;
    ~Resolve() {
      if (c.timer().enabled()) c.timer().resolve();
    }
  } resolveOnExit{ctx};

  if constexpr (Scalar<T>::kF32) {
    gemm_gpu_f32<T>(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx, fill);
  } else {
    gemm_gpu_f64<T>(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx, fill);
  }
}

template <typename T>
void gemm_gpu_fixed_workspace(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n,
                              int64_t k, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                              T beta, T* C, int64_t ldc, spla::ContextInternal& ctx,
                              void* workspace, std::size_t workspaceBytes) {
  if (m == 0 || n == 0) return;
  check_gemm_param(opA, opB, m, n, k, A, lda, B, ldb, C, ldc);
  if (opA != SPLA_OP_NONE && opA != SPLA_OP_TRANSPOSE && opA != SPLA_OP_CONJ_TRANSPOSE)
    throw spla::InvalidParameterError();
  if (opB != SPLA_OP_NONE && opB != SPLA_OP_TRANSPOSE && opB != SPLA_OP_CONJ_TRANSPOSE)
    throw spla::InvalidParameterError();
  if (m > (int64_t(1) << 30) || n > (int64_t(1) << 30)) throw spla::InvalidParameterError();
  // host memory would need device staging buffers
  if ((k > 0 && (classify(A) != MemKind::Device || classify(B) != MemKind::Device)) ||
      classify(C) != MemKind::Device)
    throw spla::GPUInvalidDevicePointerError();
  DeviceGuard guard(ctx.device());
  hip_check(hipStreamSynchronize(nullptr));
  ctx.timer().reset();
  const hipStream_t stream = ctx.compute_stream();
  std::vector<Block> keep;  // stays empty: the workspace is the caller's
  const FixedWorkspace fw{workspace, workspaceBytes};
  enqueue_device_product<T>(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx,
                            FillSpec(), stream, keep, nullptr, /*trySplit=*/false,
                            /*wholeChip=*/true, &fw);
  hip_check(hipStreamSynchronize(stream));
  if (ctx.timer().enabled()) ctx.timer().resolve();
}

template void gemm_gpu_fixed_workspace<std::complex<double>>(
    SplaOperation, SplaOperation, int64_t, int64_t, int64_t, std::complex<double>,
    const std::complex<double>*, int64_t, const std::complex<double>*, int64_t,
    std::complex<double>, std::complex<double>*, int64_t, spla::ContextInternal&, void*,
    std::size_t);

template <typename T>
void enqueue_on_device(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                       T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* C,
                       int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                       hipStream_t stream, std::vector<Block>& keep, Block* reuse,
                       bool trySplit, bool wholeChip) {
  enqueue_device_product<T>(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx, fill,
                            stream, keep, reuse, trySplit, wholeChip);
}

double fp64_estimate_seconds(spla::ContextInternal& ctx, bool complex, SplaOperation opA,
                             SplaOperation opB, int64_t m, int64_t n, int64_t k, bool wholeChip) {
  const int la = (opA == SPLA_OP_NONE) ? MNMajor : KMajor;
  const int lb = (opB == SPLA_OP_NONE) ? KMajor : MNMajor;
  return ctx.tuner().plan(complex, la, lb, m, n, k, ctx.num_cus(), wholeChip).estSeconds;
}

template void enqueue_on_device<double>(SplaOperation, SplaOperation, int64_t, int64_t, int64_t,
                                        double, const double*, int64_t, const double*, int64_t,
                                        double, double*, int64_t, spla::ContextInternal&,
                                        const FillSpec&, hipStream_t, std::vector<Block>&, Block*,
                                        bool, bool);
template void enqueue_on_device<std::complex<double>>(
    SplaOperation, SplaOperation, int64_t, int64_t, int64_t, std::complex<double>,
    const std::complex<double>*, int64_t, const std::complex<double>*, int64_t,
    std::complex<double>, std::complex<double>*, int64_t, spla::ContextInternal&, const FillSpec&,
    hipStream_t, std::vector<Block>&, Block*, bool, bool);

template void gemm_gpu<float>(SplaOperation, SplaOperation, int64_t, int64_t, int64_t, float,
                              const float*, int64_t, const float*, int64_t, float, float*, int64_t,
                              spla::ContextInternal&, FillSpec);
template void gemm_gpu<double>(SplaOperation, SplaOperation, int64_t, int64_t, int64_t, double,
                               const double*, int64_t, const double*, int64_t, double, double*,
                               int64_t, spla::ContextInternal&, FillSpec);
template void gemm_gpu<std::complex<float>>(SplaOperation, SplaOperation, int64_t, int64_t,
                                            int64_t, std::complex<float>,
                                            const std::complex<float>*, int64_t,
                                            const std::complex<float>*, int64_t,
                                            std::complex<float>, std::complex<float>*, int64_t,
                                            spla::ContextInternal&, FillSpec);
template void gemm_gpu<std::complex<double>>(SplaOperation, SplaOperation, int64_t, int64_t,
                                             int64_t, std::complex<double>,
                                             const std::complex<double>*, int64_t,
                                             const std::complex<double>*, int64_t,
                                             std::complex<double>, std::complex<double>*, int64_t,
                                             spla::ContextInternal&, FillSpec);

}  // namespace spla_amd
