// Outcome of the last FP64 / complex-FP64 product of a context.
#pragma once

namespace spla_amd {

// Values of spla_amd_ctx_set_gemm_mode (include/spla_amd/spla_amd.h).
enum GemmMode : int { GemmModeFP64 = 0, GemmModeSplitInt8 = 1, GemmModeAuto = 2 };

struct SplitReport {
  int used = 0;     // 1: the product ran on the split-precision path
  int reason = 7;   // SplitReason of the decision (0 when used)
  int moduli = 0;   // number of moduli
  int bitsA = 0;    // integer width of the scaled op(A) rows
  int bitsB = 0;    // integer width of the scaled op(B) columns
  double bound = 0; // estimated normwise error bound of the split product
};

}  // namespace spla_amd
