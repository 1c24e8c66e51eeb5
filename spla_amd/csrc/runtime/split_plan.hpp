// Host-side planning of the split-precision (Ozaki scheme II) product.
//
// Pure C++ (no HIP): builds and is tested on any host.
//
// Given per-row statistics of op(A) (rows = m) and per-column statistics of
// op(B) (columns = n) -- the largest component magnitude and the sum of
// squares over the reduction index -- the plan fixes
//   * the integer width P of the scaled operands: row i is scaled by
//     2^e_i with e_i = P - 1 - floor(log2 max_l |A_il|), so every scaled
//     component is an integer of at most P bits after rounding. The rounding
//     error bound ||dA||_F^2 <= sum_i comps * k * 2^(-2 e_i - 2) gives
//     rho_A = ||dA||_F / ||A||_F = 2^-P * S_A with S_A computed from the data,
//     so P grows with the spread of magnitudes inside the rows. The normwise
//     error of the emulated product is bounded by rho_A + rho_B + rho_A rho_B
//     (plus one rounding to FP64), and P is the smallest width that keeps this
//     below the tolerance;
//   * the number N of moduli: their product M must exceed
//     2 max_ij |(a^T b)_ij|, bounded by Cauchy-Schwarz with the scaled row norms;
//   * a fall-back to the plain FP64 kernels when the inputs are not finite, an
//     operand is zero, P would exceed 53 bits or more than maxModuli moduli
//     would be needed.
#pragma once

#include <cstdint>
#include <vector>

#include "kernels/ozaki_desc.hpp"

namespace spla_amd {

// Pairwise coprime moduli in the order they are used.
extern const int kModuliList[kMaxModuli];

struct OperandStats {
  const double* maxabs;  // per row: largest |component|; NaN when a non-finite value was seen
  const double* sumsq;   // per row: sum of squared components
  int rows;
  int64_t k;
  int comps;  // 1 real, 2 complex
};

enum SplitReason : int {
  SplitOk = 0,
  SplitNonFinite = 1,
  SplitZeroOperand = 2,
  SplitTooWide = 3,       // P > 53
  SplitTooManyModuli = 4,
  SplitRange = 5,         // norms outside the double range
  SplitNotFaster = 6,     // automatic mode: the FP64 kernel is expected to be faster
  SplitDisabled = 7,
  SplitNoWorkspace = 8    // workspace over the limit and not chunkable, or allocation failed
};

struct SplitPlan {
  bool ok = false;
  int reason = SplitDisabled;
  int PA = 0, PB = 0;
  int moduli = 0;
  double rhoA = 0, rhoB = 0, bound = 0;  // estimated normwise error bound
  double bitsNeeded = 0;
  double tolerance = 0;
  std::vector<int> expA, expB;
};

// Unit roundoff based bound of a plain FP64 product: (comps * k) * 2^-53.
double fp64_bound(int64_t k, int comps);

SplitPlan plan_split(const OperandStats& A, const OperandStats& B, double tolerance,
                     int maxModuli);

// Fewest moduli plan_split can prescribe for a tolerance, whatever the data:
// every row scale factor exceeds 1/2, so each operand needs at least
// ceil(log2(1.01 / tolerance)) bits and the moduli must cover twice that.
int min_moduli(double tolerance);

// Residue rows and reduction splits are whole multiples of two GEMM k-blocks
// (2 x 128 bytes): the GEMM pipeline runs k-blocks in pairs.
constexpr int64_t kKpadAlign = 256;
// Largest split of the reduction (bytes of int8 residues) whose INT32 sums
// cannot overflow: 128 * 128 * (2^17 - 256) < 2^31.
constexpr int64_t kMaxSplitK = (int64_t(1) << 17) - 256;
// Sustained INT8 GEMM rate of 256 x 256 tiles on a full chip, as the cost
// model assumes it (2.4-2.7e15 measured with the 16 x 16 x 64 MFMA form; the
// clock depends on the power limit).
constexpr double kInt8OpsPerSecond = 2.4e15;

// Tile shape (I8TileConfig) and reduction splits of the batched INT8 GEMM of
// one product: m x n outputs, kpad padded reduction bytes, G products.
struct I8Shape {
  int cfg = 0, bm = 0, bn = 0;
  int64_t tiles = 0, splits = 0, kSplit = 0;
  double seconds = 0;  // modelled GEMM time
};
// For each tile shape the reduction is split to stay within the INT32 range
// and to fill the chip; the shape with the shorter modelled time is taken.
I8Shape plan_i8_shape(int64_t m, int64_t n, int64_t kpad, int G, int numCUs);

void make_moduli(int count, ModuliSet& ms);
void make_crt_constants(const ModuliSet& ms, CrtConstants& cc);

}  // namespace spla_amd
