// Process-wide hierarchical wall-clock timer behind SPLA's spla_timer_* exports
// (src/timing/timing.cpp:49-86, which wrap an rt_graph::Timer).
//
// A scope is opened with start(name) and closed with stop(name) on the same
// thread; scopes opened while another is open on that thread become its
// children. Every closed scope appends its duration (seconds) to the node of
// its path, so repeated scopes at one place accumulate. Stopping a name that is
// not the innermost open scope of the thread is recorded as a warning (printed
// with the table) and otherwise ignored, as rt_graph reports unmatched pairs
// when it processes its time stamps.
//
// The library's own scopes ("pgemm_ssb_gpu", "pgemm_sbs_gpu", "gemm_gpu",
// "gemm" as in SPLA's SCOPED_TIMING sites) are recorded only while
// set_library_scopes(true) is in effect; user scopes are always recorded.
#pragma once

#include <string>

namespace spla_amd {
namespace timer {

void start(const std::string& name);
void stop(const std::string& name);
// Nested JSON object {"name": {"timings": [...], "sub-timings": {...}}, ...},
// the structure rt_graph::TimingResult::json writes.
std::string json();
// Table of every node: count, total, share of the top-level total, share of the
// parent, median, min, max (seconds), children indented below their parent.
std::string table();
void clear();

void set_library_scopes(bool enable);
bool library_scopes();

// RAII scope for the library's own timings (no-op unless library scopes are on).
class LibraryScope {
 public:
  explicit LibraryScope(const char* name) : name_(library_scopes() ? name : nullptr) {
    if (name_) start(name_);
  }
  ~LibraryScope() {
    if (name_) stop(name_);
  }
  LibraryScope(const LibraryScope&) = delete;
  LibraryScope& operator=(const LibraryScope&) = delete;

 private:
  const char* name_;
};

}  // namespace timer
}  // namespace spla_amd
