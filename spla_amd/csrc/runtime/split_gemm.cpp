#include "runtime/split_gemm.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <type_traits>

#include "kernels/launch.hpp"
#include "runtime/engine.hpp"
#include "runtime/hip_check.hpp"
#include "runtime/split_plan.hpp"
#include "runtime/timer.hpp"
#include "spla/exceptions.hpp"

namespace spla_amd {
namespace {

// longest padded residue row: offsets of the 256 rows of a tile fit 32 bits
constexpr int64_t kMaxKpad = int64_t(1) << 24;
// residue planes hold whole tiles of the INT8 GEMM (rows padded to 256)
constexpr int64_t kPlaneRowAlign = 256;
// HBM rate of the statistics and residue passes, as the automatic mode assumes
// it (5.1-5.5 TB/s measured over their reads and plane writes).
constexpr double kHbmBytesPerSecond = 5.0e12;
// statistics round trip to the host, plan and the six launches of a product
constexpr double kSplitOverheadSeconds = 150e-6;

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

template <typename T>
struct Cplx : std::false_type {};
template <>
struct Cplx<std::complex<double>> : std::true_type {};

double re_of(double a) { return a; }
double im_of(double) { return 0.0; }
double re_of(std::complex<double> a) { return a.real(); }
double im_of(std::complex<double> a) { return a.imag(); }

int stats_splits(int layout, int rows, int64_t k) {
  const int64_t blocksPerSplit = layout == KMajor ? rows : (rows + 63) / 64;
  int64_t s = (2048 + blocksPerSplit - 1) / std::max<int64_t>(1, blocksPerSplit);
  s = std::min<int64_t>(s, std::max<int64_t>(1, k / 256));
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(s, 256)));
}

}  // namespace

template <typename T>
SplitProduct<T>::SplitProduct(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n,
                              int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                              int64_t ldb, T beta, T* C, int64_t ldc,
                              spla::ContextInternal& ctx, const FillSpec& fill)
    : opA_(opA), opB_(opB), m_(m), n_(n), k_(k), alpha_(alpha), beta_(beta), A_(A), B_(B),
      C_(C), lda_(lda), ldb_(ldb), ldc_(ldc), ctx_(ctx), fill_(fill) {}

template <typename T>
bool SplitProduct<T>::enqueue_stats(hipStream_t stream) {
  constexpr bool cplx = Cplx<T>::value;
  rep_ = SplitReport();
  if (ctx_.gemm_mode() == GemmModeFP64 || k_ == 0 || m_ == 0 || n_ == 0) {
    rep_.reason = SplitDisabled;
    return false;
  }
  if (round_up(k_, kKpadAlign) > kMaxKpad) {  // residue rows address with 32-bit offsets
    rep_.reason = SplitRange;
    return false;
  }
  la_ = (opA_ == SPLA_OP_NONE) ? MNMajor : KMajor;
  lb_ = (opB_ == SPLA_OP_NONE) ? KMajor : MNMajor;
  const int mi = static_cast<int>(m_), ni = static_cast<int>(n_);
  PhaseScope ps(ctx_.timer(), PhaseSplitStats, stream);
  const int sA = stats_splits(la_, mi, k_), sB = stats_splits(lb_, ni, k_);
  const std::size_t statDoubles =
      2 * (static_cast<std::size_t>(sA) * m_ + static_cast<std::size_t>(sB) * n_) +
      2 * static_cast<std::size_t>(m_ + n_);
  dStat_ = Block(ctx_.gpu_memory(), statDoubles * sizeof(double));
  double* partA = dStat_.get<double>();
  double* partB = partA + 2 * static_cast<std::size_t>(sA) * m_;
  double* outA = partB + 2 * static_cast<std::size_t>(sB) * n_;  // [max m | sum m | max n | sum n]
  hip_check(launch_split_stats(cplx, la_, reinterpret_cast<const double*>(A_), lda_, mi, k_,
                               partA, sA, outA, outA + m_, stream));
  hip_check(launch_split_stats(cplx, lb_, reinterpret_cast<const double*>(B_), ldb_, ni, k_,
                               partB, sB, outA + 2 * m_, outA + 2 * m_ + n_, stream));
  hStat_ = Block(ctx_.pinned_memory(), 2 * static_cast<std::size_t>(m_ + n_) * sizeof(double));
  hip_check(hipMemcpyAsync(hStat_.get(), outA, hStat_.bytes(), hipMemcpyDeviceToHost, stream));
  return true;
}

template <typename T>
double SplitProduct<T>::lower_bound_seconds() const {
  constexpr bool cplx = Cplx<T>::value;
  constexpr int comps = cplx ? 2 : 1;
  constexpr int products = cplx ? 3 : 1;
  if (m_ == 0 || n_ == 0 || k_ == 0) return 0.0;
  const double tol = std::min(ctx_.split_tolerance(), std::max(fp64_bound(k_, comps), 1e-300));
  const double G = static_cast<double>(min_moduli(tol)) * products;
  const double kpad = static_cast<double>(round_up(k_, kKpadAlign));
  const double md = static_cast<double>(m_), nd = static_cast<double>(n_);
  // the cost plan() estimates, with no tile padding, a full chip and one split
  const double ops = 2.0 * G * md * nd * kpad;
  const double bytes = (md + nd) * static_cast<double>(k_) * comps * 8.0 * 2 + G * (md + nd) * kpad +
                       G * md * nd;
  return ops / kInt8OpsPerSecond + bytes / kHbmBytesPerSecond + kSplitOverheadSeconds;
}

template <typename T>
bool SplitProduct<T>::may_beat_fp64(double fp64Seconds) {
  if (m_ == 0 || n_ == 0 || k_ == 0) return true;
  if (lower_bound_seconds() < 0.8 * fp64Seconds) return true;
  rep_.reason = SplitNotFaster;
  return false;
}

template <typename T>
bool SplitProduct<T>::plan(double fp64Seconds) {
  constexpr bool cplx = Cplx<T>::value;
  constexpr int comps = cplx ? 2 : 1;
  constexpr int products = cplx ? 3 : 1;
  const int mi = static_cast<int>(m_), ni = static_cast<int>(n_);
  const double* h = hStat_.get<double>();
  OperandStats stA{h, h + m_, mi, k_, comps};
  OperandStats stB{h + 2 * m_, h + 2 * m_ + n_, ni, k_, comps};
  const double tol = std::min(ctx_.split_tolerance(), std::max(fp64_bound(k_, comps), 1e-300));
  plan_ = plan_split(stA, stB, tol, std::min(ctx_.split_max_moduli(), kMaxModuli));
  dStat_.reset();
  hStat_.reset();
  rep_.reason = plan_.reason;
  rep_.moduli = plan_.moduli;
  rep_.bitsA = plan_.PA;
  rep_.bitsB = plan_.PB;
  rep_.bound = plan_.bound;
  if (!plan_.ok) return false;

  G_ = plan_.moduli * products;
  kpad_ = round_up(k_, kKpadAlign);
  mpad_ = round_up(m_, kPlaneRowAlign);
  npad_ = round_up(n_, kPlaneRowAlign);
  // INT8 GEMM tile shape and reduction splits
  const int numCUs = ctx_.num_cus();
  const I8Shape sh = plan_i8_shape(m_, n_, kpad_, static_cast<int>(G_), numCUs);
  cfg_ = sh.cfg;
  bm_ = sh.bm;
  bn_ = sh.bn;
  splits_ = sh.splits;
  kSplit_ = sh.kSplit;
  const int64_t tiles = sh.tiles;
  if (ctx_.gemm_mode() == GemmModeAuto) {
    // INT8 work on whole tiles at the rate of the chosen tile, scaled by how
    // well the workgroups fill the chip (last partial wave), plus the passes
    // over A, B and C and the host plan between them
    const double blocks = static_cast<double>(tiles) * G_ * splits_;
    const double waves = std::ceil(blocks / numCUs);
    const double fill = blocks / (waves * numCUs);
    const double ops = 2.0 * G_ * static_cast<double>(((m_ + bm_ - 1) / bm_) * bm_) *
                       static_cast<double>(((n_ + bn_ - 1) / bn_) * bn_) * kpad_;
    const double rate = kInt8OpsPerSecond * (cfg_ == I8TileL ? 1.0 : 0.58) * fill;
    const double bytes = static_cast<double>(m_ + n_) * k_ * comps * 8.0 * 2 +
                         static_cast<double>(G_) * (m_ + n_) * kpad_ +
                         static_cast<double>(G_) * m_ * n_ * (splits_ == 1 ? 1.0 : 4.0 * splits_);
    const double est = ops / rate + bytes / kHbmBytesPerSecond + kSplitOverheadSeconds;
    if (!(est < 0.8 * fp64Seconds)) {
      rep_.reason = SplitNotFaster;
      return false;
    }
  }
  return true;
}

template <typename T>
std::size_t SplitProduct<T>::workspace_bytes() const {
  const std::size_t planes = round_up(static_cast<int64_t>(G_) * mpad_ * kpad_, 256) +
                             round_up(static_cast<int64_t>(G_) * npad_ * kpad_, 256);
  const std::size_t outPlane = static_cast<std::size_t>(mpad_) * n_;
  const std::size_t out = splits_ == 1 ? G_ * outPlane : splits_ * G_ * outPlane * 4;
  return planes + out;
}

template <typename T>
void SplitProduct<T>::enqueue_rest(hipStream_t stream, void* workspace, std::vector<Block>& keep) {
  constexpr bool cplx = Cplx<T>::value;
  constexpr int products = cplx ? 3 : 1;
  const int mi = static_cast<int>(m_), ni = static_cast<int>(n_);
  PhaseTimer& timer = ctx_.timer();

  // exponents and CRT constants to the device
  ModuliSet ms;
  make_moduli(plan_.moduli, ms);
  CrtConstants cc;
  make_crt_constants(ms, cc);
  const std::size_t expBytes = round_up(static_cast<int64_t>((m_ + n_) * sizeof(int)), 256);
  const std::size_t upBytes = expBytes + sizeof(CrtConstants);
  keep.emplace_back(ctx_.pinned_memory(), upBytes);
  char* hUp = keep.back().get<char>();
  std::memcpy(hUp, plan_.expA.data(), m_ * sizeof(int));
  std::memcpy(hUp + m_ * sizeof(int), plan_.expB.data(), n_ * sizeof(int));
  std::memcpy(hUp + expBytes, &cc, sizeof(cc));
  keep.emplace_back(ctx_.gpu_memory(), upBytes);
  char* dUp = keep.back().get<char>();
  hip_check(hipMemcpyAsync(dUp, hUp, upBytes, hipMemcpyHostToDevice, stream));
  const int* dExpA = reinterpret_cast<const int*>(dUp);
  const int* dExpB = dExpA + m_;
  const CrtConstants* dCC = reinterpret_cast<const CrtConstants*>(dUp + expBytes);

  // workspace: residue planes of A and B, INT8 GEMM output
  if (!workspace) {
    keep.emplace_back(ctx_.gpu_memory(), workspace_bytes());
    workspace = keep.back().get();
  }
  int8_t* RA = static_cast<int8_t*>(workspace);
  int8_t* RB = RA + round_up(static_cast<int64_t>(G_) * mpad_ * kpad_, 256);
  void* out = RB + round_up(static_cast<int64_t>(G_) * npad_ * kpad_, 256);

  // residue planes
  ResidueDesc rd;
  std::memset(&rd, 0, sizeof(rd));
  rd.X = reinterpret_cast<const double*>(A_);
  rd.ld = lda_;
  rd.rows = mi;
  rd.k = k_;
  rd.kpad = kpad_;
  rd.expo = dExpA;
  rd.imagSign = (cplx && opA_ == SPLA_OP_CONJ_TRANSPOSE) ? -1 : 1;
  rd.R = RA;
  rd.plane = mpad_ * kpad_;
  rd.ms = ms;
  {
    PhaseScope ps(timer, PhaseSplitResidues, stream);
    hip_check(launch_residues(cplx, la_, rd, stream));
    rd.X = reinterpret_cast<const double*>(B_);
    rd.ld = ldb_;
    rd.rows = ni;
    rd.expo = dExpB;
    rd.imagSign = (cplx && opB_ == SPLA_OP_CONJ_TRANSPOSE) ? -1 : 1;
    rd.R = RB;
    rd.plane = npad_ * kpad_;
    hip_check(launch_residues(cplx, lb_, rd, stream));
  }

  // batched INT8 GEMM
  I8GemmDesc gd;
  std::memset(&gd, 0, sizeof(gd));
  gd.RA = RA;
  gd.RB = RB;
  gd.planeA = mpad_ * kpad_;
  gd.planeB = npad_ * kpad_;
  gd.kpad = kpad_;
  gd.m = mi;
  gd.n = ni;
  gd.G = G_;
  gd.products = products;
  gd.splits = static_cast<int>(splits_);
  gd.kSplit = kSplit_;
  gd.tilesM = static_cast<int>((m_ + bm_ - 1) / bm_);
  gd.tilesN = static_cast<int>((n_ + bn_ - 1) / bn_);
  gd.ldo = mpad_;  // covers every tile row; a multiple of 4 for word stores
  gd.outPlane = gd.ldo * n_;
  gd.outSplitStride = static_cast<int64_t>(G_) * gd.outPlane;
  gd.ms = ms;
  gd.fillMode = fill_.mode;
  gd.fillRowOff = fill_.rowOff;
  gd.fillColOff = fill_.colOff;
  gd.out = out;  // tiles skipped by the fill mode are neither written nor read
  {
    PhaseScope ps(timer, PhaseSplitInt8Gemm, stream);
    hip_check(launch_i8_gemm(gd, cfg_, stream));
  }

  // reconstruction and alpha / beta update
  CrtDesc cd;
  std::memset(&cd, 0, sizeof(cd));
  cd.W = out;
  cd.in8 = splits_ == 1;
  cd.splits = static_cast<int>(splits_);
  cd.G = G_;
  cd.products = products;
  cd.ldo = gd.ldo;
  cd.plane = gd.outPlane;
  cd.m = mi;
  cd.n = ni;
  cd.expA = dExpA;
  cd.expB = dExpB;
  cd.cc = dCC;
  cd.ms = ms;
  cd.C = reinterpret_cast<double*>(C_);
  cd.ldc = ldc_;
  cd.alphaRe = re_of(alpha_);
  cd.alphaIm = im_of(alpha_);
  cd.betaRe = re_of(beta_);
  cd.betaIm = im_of(beta_);
  cd.betaZero = (cd.betaRe == 0.0 && cd.betaIm == 0.0) ? 1 : 0;
  cd.fillMode = fill_.mode;
  cd.fillRowOff = fill_.rowOff;
  cd.fillColOff = fill_.colOff;
  cd.tileRows = bm_;
  cd.tileCols = bn_;
  {
    PhaseScope ps(timer, PhaseSplitCrt, stream);
    hip_check(launch_crt(cplx, cd, cc.limbs, stream));
  }
  rep_.used = 1;
}

template <typename T>
int64_t SplitProduct<T>::max_chunk_k(std::size_t limit) const {
  // the planes grow with the reduction length, the GEMM output does not
  const std::size_t planes = round_up(static_cast<int64_t>(G_) * mpad_ * kpad_, 256) +
                             round_up(static_cast<int64_t>(G_) * npad_ * kpad_, 256);
  const std::size_t out = workspace_bytes() - planes;
  if (limit <= out + 512) return 0;
  const int64_t perK = static_cast<int64_t>(G_) * (mpad_ + npad_);
  return (static_cast<int64_t>(limit - out - 512) / perK) / kKpadAlign * kKpadAlign;
}

namespace {

// Device workspace of at least `bytes`: the caller's reusable block (grown when
// too small; the stream has drained, so the old one may go back to the pool) or
// a block appended to `keep`. nullptr when the device memory is exhausted.
void* split_workspace(spla::ContextInternal& ctx, std::size_t bytes, Block* reuse,
                      std::vector<Block>& keep) {
  try {
    if (reuse) {
      if (reuse->bytes() < bytes) {
        reuse->reset();
        *reuse = Block(ctx.gpu_memory(), bytes);
      }
      return reuse->get();
    }
    keep.emplace_back(ctx.gpu_memory(), bytes);
    return keep.back().get();
  } catch (const spla::GPUAllocationError&) {
    return nullptr;
  }
}

// Shortest reduction chunk worth a split product of its own.
constexpr int64_t kMinChunkK = 4096;

}  // namespace

template <typename T>
bool enqueue_split_product(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                           T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                           T* C, int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                           double fp64Seconds, hipStream_t stream, std::vector<Block>& keep,
                           Block* reuse) {
  SplitProduct<T> p(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx, fill);
  bool ok = (ctx.gemm_mode() != GemmModeAuto || p.may_beat_fp64(fp64Seconds)) &&
            p.enqueue_stats(stream);
  if (ok) {
    timer::LibraryScope scope("split_plan");  // statistics wait and the host plan
    hip_check(hipStreamSynchronize(stream));
    ok = p.plan(fp64Seconds);
  }
  if (!ok) {
    ctx.last_split() = p.report();
    return false;
  }
  const std::size_t limit = ctx.split_workspace_limit();
  if (p.workspace_bytes() <= limit) {
    void* ws = split_workspace(ctx, p.workspace_bytes(), reuse, keep);
    if (!ws) {
      SplitReport r = p.report();
      r.used = 0;
      r.reason = SplitNoWorkspace;
      ctx.last_split() = r;
      return false;
    }
    p.enqueue_rest(stream, ws, keep);
    ctx.last_split() = p.report();
    return true;
  }

  // Over the workspace limit: the reduction runs in chunks, each a split product
  // of its own (statistics, plan, residues, INT8 GEMM, reconstruction) that adds
  // into C in FP64, beta on the first chunk only. Each chunk meets the tolerance
  // relative to its own ||A_c|| ||B_c||, and sum_c ||A_c|| ||B_c|| <= ||A|| ||B||.
  const int64_t kmax = p.max_chunk_k(limit);
  if (kmax < kMinChunkK) {
    SplitReport r = p.report();
    r.used = 0;
    r.reason = SplitNoWorkspace;
    ctx.last_split() = r;
    return false;
  }
  const int64_t chunks = (k + kmax - 1) / kmax;
  const int64_t kc = std::min(kmax, round_up((k + chunks - 1) / chunks, kKpadAlign));
  Block own;
  Block* ws = reuse ? reuse : &own;
  SplitReport agg;
  agg.used = 1;
  agg.reason = SplitOk;
  for (int64_t k0 = 0; k0 < k; k0 += kc) {
    const int64_t kcur = std::min(kc, k - k0);
    const T* cA = opA == SPLA_OP_NONE ? A + k0 * lda : A + k0;
    const T* cB = opB == SPLA_OP_NONE ? B + k0 : B + k0 * ldb;
    const T cBeta = k0 == 0 ? beta : T(1);
    SplitProduct<T> q(opA, opB, m, n, kcur, alpha, cA, lda, cB, ldb, cBeta, C, ldc, ctx, fill);
    bool split = q.enqueue_stats(stream);
    if (split) {
      timer::LibraryScope scope("split_plan");
      hip_check(hipStreamSynchronize(stream));  // also frees the previous chunk's workspace
      // the whole product was found worth splitting; chunks are not re-judged
      split = q.plan(std::numeric_limits<double>::infinity()) && q.workspace_bytes() <= limit;
    }
    void* w = split ? split_workspace(ctx, q.workspace_bytes(), ws, keep) : nullptr;
    if (w) {
      q.enqueue_rest(stream, w, keep);
      agg.moduli = std::max(agg.moduli, q.report().moduli);
      agg.bitsA = std::max(agg.bitsA, q.report().bitsA);
      agg.bitsB = std::max(agg.bitsB, q.report().bitsB);
      agg.bound = std::max(agg.bound, q.report().bound);
    } else {
      // this chunk on the FP64 kernels (same beta rule)
      enqueue_on_device<T>(opA, opB, m, n, kcur, alpha, cA, lda, cB, ldb, cBeta, C, ldc, ctx,
                           fill, stream, keep, nullptr, /*trySplit=*/false, /*wholeChip=*/true);
      agg.used = 0;
      agg.reason = q.report().reason != SplitOk ? q.report().reason : SplitNoWorkspace;
    }
  }
  if (!reuse) keep.push_back(std::move(own));
  ctx.last_split() = agg;
  return true;
}

template class SplitProduct<double>;
template class SplitProduct<std::complex<double>>;

template bool enqueue_split_product<double>(SplaOperation, SplaOperation, int64_t, int64_t,
                                            int64_t, double, const double*, int64_t,
                                            const double*, int64_t, double, double*, int64_t,
                                            spla::ContextInternal&, const FillSpec&, double,
                                            hipStream_t, std::vector<Block>&, Block*);
template bool enqueue_split_product<std::complex<double>>(
    SplaOperation, SplaOperation, int64_t, int64_t, int64_t, std::complex<double>,
    const std::complex<double>*, int64_t, const std::complex<double>*, int64_t,
    std::complex<double>, std::complex<double>*, int64_t, spla::ContextInternal&, const FillSpec&,
    double, hipStream_t, std::vector<Block>&, Block*);

}  // namespace spla_amd
