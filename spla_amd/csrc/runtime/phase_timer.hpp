// Optional per-phase kernel timing of a context's products: HIP events are
// recorded on the launch stream around every phase and resolved after the
// stream has drained. Times of a phase add up over the launches of one call.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <vector>

namespace spla_amd {

enum Phase : int {
  PhaseGemmF64 = 0,    // FP64 MFMA GEMM kernels
  PhaseSplitKReduce,   // split-K reduction and alpha / beta update
  PhaseSplitStats,     // split precision: operand statistics (+ host plan)
  PhaseSplitResidues,  // split precision: residue planes
  PhaseSplitInt8Gemm,  // split precision: batched INT8 GEMM
  PhaseSplitCrt,       // split precision: reconstruction and update
  NumPhases
};

class PhaseTimer {
public:
  ~PhaseTimer();
  bool enabled() const { return enabled_; }
  void set_enabled(bool e) { enabled_ = e; }
  void reset();                       // start of an API call
  void begin(int phase, hipStream_t s);
  void end(hipStream_t s);
  void resolve();                     // after the stream has drained
  const std::array<double, NumPhases>& ms() const { return ms_; }

private:
  hipEvent_t event();
  bool enabled_ = false;
  std::vector<hipEvent_t> pool_;
  std::size_t used_ = 0;
  struct Span {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Span> spans_;
  int open_ = -1;
  hipEvent_t openEvent_ = nullptr;
  std::array<double, NumPhases> ms_{};
};

// Scoped phase: records only when the timer is enabled.
class PhaseScope {
public:
  PhaseScope(PhaseTimer& t, int phase, hipStream_t s) : t_(t), s_(s), on_(t.enabled()) {
    if (on_) t_.begin(phase, s_);
  }
  ~PhaseScope() {
    if (on_) t_.end(s_);
  }

private:
  PhaseTimer& t_;
  hipStream_t s_;
  bool on_;
};

}  // namespace spla_amd
