#include "runtime/split_gemm.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <type_traits>

#include "kernels/launch.hpp"
#include "runtime/engine.hpp"
#include "runtime/hip_check.hpp"
#include "runtime/split_plan.hpp"

namespace spla_amd {
namespace {

// Residue rows and reduction splits are whole multiples of two GEMM k-blocks
// (2 x 128 bytes): the GEMM pipeline runs k-blocks in pairs.
constexpr int64_t kKpadAlign = 256;
// Largest split of the reduction (bytes of int8 residues) whose INT32 sums
// cannot overflow: 128 * 128 * (2^17 - 256) < 2^31.
constexpr int64_t kMaxSplitK = (int64_t(1) << 17) - 256;
// longest padded residue row: offsets of the 256 rows of a tile fit 32 bits
constexpr int64_t kMaxKpad = int64_t(1) << 24;
// residue planes hold whole tiles of the INT8 GEMM (rows padded to 256)
constexpr int64_t kPlaneRowAlign = 256;
// Sustained rates assumed by the automatic mode.
constexpr double kInt8OpsPerSecond = 2.0e15;
constexpr double kHbmBytesPerSecond = 4.0e12;

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
bool enqueue_split_product(SplaOperation opA, SplaOperation opB, int64_t m, int64_t n, int64_t k,
                           T alpha, const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                           T* C, int64_t ldc, spla::ContextInternal& ctx, const FillSpec& fill,
                           double fp64Seconds, hipStream_t stream, std::vector<Block>& keep) {
  constexpr bool cplx = Cplx<T>::value;
  constexpr int comps = cplx ? 2 : 1;
  constexpr int products = cplx ? 3 : 1;
  SplitReport& rep = ctx.last_split();
  rep = SplitReport();
  const int mode = ctx.gemm_mode();
  if (mode == GemmModeFP64 || k == 0 || m == 0 || n == 0) {
    rep.reason = SplitDisabled;
    return false;
  }
  if (round_up(k, kKpadAlign) > kMaxKpad) {  // residue rows address with 32-bit offsets
    rep.reason = SplitRange;
    return false;
  }
  const int la = (opA == SPLA_OP_NONE) ? MNMajor : KMajor;
  const int lb = (opB == SPLA_OP_NONE) ? KMajor : MNMajor;
  const int mi = static_cast<int>(m), ni = static_cast<int>(n);

  // 1. operand statistics
  PhaseTimer& timer = ctx.timer();
  const bool timing = timer.enabled();
  if (timing) timer.begin(PhaseSplitStats, stream);
  const int sA = stats_splits(la, mi, k), sB = stats_splits(lb, ni, k);
  const std::size_t statDoubles = 2 * (static_cast<std::size_t>(sA) * m + static_cast<std::size_t>(sB) * n) +
                                  2 * static_cast<std::size_t>(m + n);
  keep.emplace_back(ctx.gpu_memory(), statDoubles * sizeof(double));
  double* dStat = keep.back().get<double>();
  double* partA = dStat;
  double* partB = partA + 2 * static_cast<std::size_t>(sA) * m;
  double* outA = partB + 2 * static_cast<std::size_t>(sB) * n;  // [max m | sum m | max n | sum n]
  hip_check(launch_split_stats(cplx, la, reinterpret_cast<const double*>(A), lda, mi, k, partA,
                               sA, outA, outA + m, stream));
  hip_check(launch_split_stats(cplx, lb, reinterpret_cast<const double*>(B), ldb, ni, k, partB,
                               sB, outA + 2 * m, outA + 2 * m + n, stream));
  Block hStat(ctx.pinned_memory(), 2 * static_cast<std::size_t>(m + n) * sizeof(double));
  hip_check(hipMemcpyAsync(hStat.get(), outA, hStat.bytes(), hipMemcpyDeviceToHost, stream));
  if (timing) timer.end(stream);
  hip_check(hipStreamSynchronize(stream));
  const double* h = hStat.get<double>();

  // 2. plan on the host
  OperandStats stA{h, h + m, mi, k, comps};
  OperandStats stB{h + 2 * m, h + 2 * m + n, ni, k, comps};
  const double tol = std::min(ctx.split_tolerance(), std::max(fp64_bound(k, comps), 1e-300));
  SplitPlan plan = plan_split(stA, stB, tol, std::min(ctx.split_max_moduli(), kMaxModuli));
  rep.reason = plan.reason;
  rep.moduli = plan.moduli;
  rep.bitsA = plan.PA;
  rep.bitsB = plan.PB;
  rep.bound = plan.bound;
  if (!plan.ok) return false;

  const int G = plan.moduli * products;
  const int64_t kpad = round_up(k, kKpadAlign);
  if (mode == GemmModeAuto) {
    const double ops = 2.0 * G * static_cast<double>(m) * n * kpad;
    const double bytes = static_cast<double>(m + n) * k * comps * 8.0 * 2 +
                         static_cast<double>(G) * (m + n) * kpad +
                         static_cast<double>(G) * m * n * 4.0;
    const double est = ops / kInt8OpsPerSecond + bytes / kHbmBytesPerSecond + 60e-6;
    if (!(est < 0.8 * fp64Seconds)) {
      rep.reason = SplitNotFaster;
      return false;
    }
  }

  // 3. exponents and CRT constants to the device
  ModuliSet ms;
  make_moduli(plan.moduli, ms);
  CrtConstants cc;
  make_crt_constants(ms, cc);
  const std::size_t expBytes = round_up(static_cast<int64_t>((m + n) * sizeof(int)), 256);
  const std::size_t upBytes = expBytes + sizeof(CrtConstants);
  keep.emplace_back(ctx.pinned_memory(), upBytes);
  char* hUp = keep.back().get<char>();
  std::memcpy(hUp, plan.expA.data(), m * sizeof(int));
  std::memcpy(hUp + m * sizeof(int), plan.expB.data(), n * sizeof(int));
  std::memcpy(hUp + expBytes, &cc, sizeof(cc));
  keep.emplace_back(ctx.gpu_memory(), upBytes);
  char* dUp = keep.back().get<char>();
  hip_check(hipMemcpyAsync(dUp, hUp, upBytes, hipMemcpyHostToDevice, stream));
  const int* dExpA = reinterpret_cast<const int*>(dUp);
  const int* dExpB = dExpA + m;
  const CrtConstants* dCC = reinterpret_cast<const CrtConstants*>(dUp + expBytes);

  // 4. residue planes
  const int64_t mpad = round_up(m, kPlaneRowAlign), npad = round_up(n, kPlaneRowAlign);
  keep.emplace_back(ctx.gpu_memory(), static_cast<std::size_t>(G) * mpad * kpad);
  int8_t* RA = keep.back().get<int8_t>();
  keep.emplace_back(ctx.gpu_memory(), static_cast<std::size_t>(G) * npad * kpad);
  int8_t* RB = keep.back().get<int8_t>();
  ResidueDesc rd;
  std::memset(&rd, 0, sizeof(rd));
  rd.X = reinterpret_cast<const double*>(A);
  rd.ld = lda;
  rd.rows = mi;
  rd.k = k;
  rd.kpad = kpad;
  rd.expo = dExpA;
  rd.imagSign = (cplx && opA == SPLA_OP_CONJ_TRANSPOSE) ? -1 : 1;
  rd.R = RA;
  rd.plane = mpad * kpad;
  rd.ms = ms;
  if (timing) timer.begin(PhaseSplitResidues, stream);
  hip_check(launch_residues(cplx, la, rd, stream));
  rd.X = reinterpret_cast<const double*>(B);
  rd.ld = ldb;
  rd.rows = ni;
  rd.expo = dExpB;
  rd.imagSign = (cplx && opB == SPLA_OP_CONJ_TRANSPOSE) ? -1 : 1;
  rd.R = RB;
  rd.plane = npad * kpad;
  hip_check(launch_residues(cplx, lb, rd, stream));
  if (timing) timer.end(stream);

  // 5. batched INT8 GEMM
  const int numCUs = ctx.num_cus();
  int cfg = I8TileL;
  int bm, bn;
  i8_tile_shape(cfg, bm, bn);
  int64_t tiles = ((m + bm - 1) / bm) * ((n + bn - 1) / bn);
  if (tiles * G < 2 * numCUs) {
    cfg = I8TileM;
    i8_tile_shape(cfg, bm, bn);
    tiles = ((m + bm - 1) / bm) * ((n + bn - 1) / bn);
  }
  int64_t splits = (kpad + kMaxSplitK - 1) / kMaxSplitK;
  while (tiles * G * splits < 2 * numCUs && kpad / (splits + 1) >= 2048) ++splits;
  const int64_t kSplit = round_up((kpad + splits - 1) / splits, kKpadAlign);
  splits = (kpad + kSplit - 1) / kSplit;
  I8GemmDesc gd;
  std::memset(&gd, 0, sizeof(gd));
  gd.RA = RA;
  gd.RB = RB;
  gd.planeA = mpad * kpad;
  gd.planeB = npad * kpad;
  gd.kpad = kpad;
  gd.m = mi;
  gd.n = ni;
  gd.G = G;
  gd.products = products;
  gd.splits = static_cast<int>(splits);
  gd.kSplit = kSplit;
  gd.tilesM = static_cast<int>((m + bm - 1) / bm);
  gd.tilesN = static_cast<int>((n + bn - 1) / bn);
  gd.ldo = mpad;  // covers every tile row; a multiple of 4 for word stores
  gd.outPlane = gd.ldo * n;
  gd.outSplitStride = static_cast<int64_t>(G) * gd.outPlane;
  gd.ms = ms;
  gd.fillMode = fill.mode;
  gd.fillRowOff = fill.rowOff;
  gd.fillColOff = fill.colOff;
  const std::size_t outBytes = splits == 1
                                   ? static_cast<std::size_t>(G) * gd.outPlane
                                   : static_cast<std::size_t>(splits) * G * gd.outPlane * 4;
  keep.emplace_back(ctx.gpu_memory(), outBytes);
  gd.out = keep.back().get();  // tiles skipped by the fill mode are neither written nor read
  {
    PhaseScope ps(timer, PhaseSplitInt8Gemm, stream);
    hip_check(launch_i8_gemm(gd, cfg, stream));
  }

  // 6. reconstruction and alpha / beta update
  CrtDesc cd;
  std::memset(&cd, 0, sizeof(cd));
  cd.W = gd.out;
  cd.in8 = splits == 1;
  cd.splits = static_cast<int>(splits);
  cd.G = G;
  cd.products = products;
  cd.ldo = gd.ldo;
  cd.plane = gd.outPlane;
  cd.m = mi;
  cd.n = ni;
  cd.expA = dExpA;
  cd.expB = dExpB;
  cd.cc = dCC;
  cd.ms = ms;
  cd.C = reinterpret_cast<double*>(C);
  cd.ldc = ldc;
  cd.alphaRe = re_of(alpha);
  cd.alphaIm = im_of(alpha);
  cd.betaRe = re_of(beta);
  cd.betaIm = im_of(beta);
  cd.betaZero = (cd.betaRe == 0.0 && cd.betaIm == 0.0) ? 1 : 0;
  cd.fillMode = fill.mode;
  cd.fillRowOff = fill.rowOff;
  cd.fillColOff = fill.colOff;
  cd.tileRows = bm;
  cd.tileCols = bn;
  {
    PhaseScope ps(timer, PhaseSplitCrt, stream);
    hip_check(launch_crt(cplx, cd, cc.limbs, stream));
  }
  rep.used = 1;
  return true;
}

template bool enqueue_split_product<double>(SplaOperation, SplaOperation, int64_t, int64_t,
                                            int64_t, double, const double*, int64_t,
                                            const double*, int64_t, double, double*, int64_t,
                                            spla::ContextInternal&, const FillSpec&, double,
                                            hipStream_t, std::vector<Block>&);
template bool enqueue_split_product<std::complex<double>>(
    SplaOperation, SplaOperation, int64_t, int64_t, int64_t, std::complex<double>,
    const std::complex<double>*, int64_t, const std::complex<double>*, int64_t,
    std::complex<double>, std::complex<double>*, int64_t, spla::ContextInternal&, const FillSpec&,
    double, hipStream_t, std::vector<Block>&);

}  // namespace spla_amd
