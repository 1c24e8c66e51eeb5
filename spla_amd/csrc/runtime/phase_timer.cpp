#include "runtime/phase_timer.hpp"

#include "runtime/hip_check.hpp"

namespace spla_amd {

PhaseTimer::~PhaseTimer() {
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t PhaseTimer::event() {
  if (used_ == pool_.size()) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e));
    pool_.push_back(e);
  }
  return pool_[used_++];
}

void PhaseTimer::reset() {
  used_ = 0;
  spans_.clear();
  open_ = -1;
  ms_.fill(0.0);
}

void PhaseTimer::begin(int phase, hipStream_t s) {
  openEvent_ = event();
  hip_check(hipEventRecord(openEvent_, s));
  open_ = phase;
}

void PhaseTimer::end(hipStream_t s) {
  if (open_ < 0) return;
  hipEvent_t b = event();
  hip_check(hipEventRecord(b, s));
  spans_.push_back({open_, openEvent_, b});
  open_ = -1;
}

void PhaseTimer::resolve() {
  ms_.fill(0.0);
  for (const Span& sp : spans_) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, sp.a, sp.b) == hipSuccess) ms_[sp.phase] += t;
    else (void)hipGetLastError();
  }
}

}  // namespace spla_amd
