// Host/device shared descriptors of the split-precision (Ozaki scheme II)
// product kernels.
//
// The product op(A) op(B) is emulated exactly on integers: every row i of
// op(A) is scaled by 2^ea[i] and every column j of op(B) by 2^eb[j] and rounded
// to integers a, b with |a|, |b| <= 2^P. The integer product a^T b is formed
// modulo pairwise coprime moduli m_t <= 256 with one INT8 MFMA GEMM per modulus
// (INT32 accumulation, symmetric residues in [-128, 127]), reconstructed by the
// Chinese remainder theorem (exact FP64 limb sums, 192-bit integer assembly),
// converted to FP64 and scaled back by 2^-(ea[i] + eb[j]). Complex products use three integer
// products per modulus (T1 = Re a Re b, T2 = Im a Im b,
// T3 = (Re a + Im a)(Re b + Im b); Re = T1 - T2, Im = T3 - T1 - T2), which is
// exact in modular arithmetic.
#pragma once

#include <cstdint>

#include "kernels/residue_arith.hpp"

namespace spla_amd {

constexpr int kMaxModuli = 20;

struct ModuliSet {
  int count;
  int m[kMaxModuli];
  double inv[kMaxModuli];  // 1.0 / m
  // constants of the FP32 residue arithmetic (residue_arith.hpp)
  residue::ModConst kc[kMaxModuli];
};

// CRT constants of one moduli set (little-endian 64-bit limbs).
constexpr int kCrtLimbBits = 40;
constexpr int kMaxCrtLimbs = 4;  // 160 bits cover 20 moduli of at most 8 bits

struct CrtConstants {
  uint64_t W[kMaxModuli][3];  // W_t = (M / m_t) * ((M / m_t)^-1 mod m_t), W_t < M
  uint64_t M[3];              // product of the moduli
  uint64_t halfM[3];          // floor(M / 2)
  uint64_t lowM[3];           // M - 1 - floor(M / 2)
  double Md;                  // M rounded to double
  // 40-bit limbs (crt_arith.hpp): W_t = sum_j Wl[t][j] 2^(40 j), M likewise
  int limbs;                  // L = ceil(bits(M) / 40)
  double Wl[kMaxModuli][kMaxCrtLimbs];
  double Ml[kMaxCrtLimbs];
  double cq[kMaxCrtLimbs];    // 2^(40 j) / M
};

// Operand read as `rows` rows of `k` reduction elements.
struct ResidueDesc {
  const double* X;  // operand origin (doubles)
  int64_t ld;       // leading dimension in elements
  int rows;
  int64_t k;
  int64_t kpad;     // padded reduction length of the residue planes (multiple of 256)
  const int* expo;  // per-row scaling exponent
  int imagSign;     // -1: the operand is read conjugated (complex only)
  int8_t* R;        // planes [count * products][rows][kpad]
  int64_t plane;    // rows * kpad
  ModuliSet ms;
};

struct I8GemmDesc {
  const int8_t* RA;  // planes [G][m][kpad]
  const int8_t* RB;  // planes [G][n][kpad]
  int64_t planeA, planeB;
  int64_t kpad;
  int m, n;
  int G;         // number of products (moduli * products per modulus)
  int products;  // 1 (real) or 3 (complex)
  int splits;
  int64_t kSplit;  // reduction bytes per split (multiple of 256)
  int tilesM, tilesN;
  // uint8 biased residues r + 128 (r in [-128, 127]) [G][n][ldo] (splits == 1)
  // or int32 sums [splits][G'][n][ldo]
  void* out;
  int64_t ldo;       // rows of an output column: a multiple of 4 that covers every tile row
  int64_t outPlane;  // ldo * n
  int64_t outSplitStride;  // int32 sums: elements from one split to the next (G' ldo n)
  ModuliSet ms;
  int fillMode;
  int64_t fillRowOff, fillColOff;
};

struct CrtDesc {
  const void* W;  // GEMM output (see I8GemmDesc::out)
  int in8;        // 1: uint8 biased residues, 0: int32 split sums
  int splits;
  int G;
  int products;
  int64_t ldo;    // rows of an output column (multiple of 4)
  int64_t plane;  // ldo * n
  int m, n;
  const int* expA;
  const int* expB;
  const CrtConstants* cc;  // device copy
  ModuliSet ms;
  double* C;
  int64_t ldc;
  double alphaRe, alphaIm, betaRe, betaIm;
  int betaZero;
  int fillMode;
  int64_t fillRowOff, fillColOff;
  int tileRows, tileCols;  // tile grid of the INT8 GEMM (for the fill predicate)
};

// Tile configurations of the INT8 GEMM.
enum I8TileConfig : int {
  I8TileL = 0,  // 256 x 256 block, 4 waves of 128 x 128
  I8TileM = 1,  // 128 x 128 block, 4 waves of 64 x 64
};

inline void i8_tile_shape(int cfg, int& bm, int& bn) {
  if (cfg == I8TileM) {
    bm = bn = 128;
  } else {
    bm = bn = 256;
  }
}

}  // namespace spla_amd
