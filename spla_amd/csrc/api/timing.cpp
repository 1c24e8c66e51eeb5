// spla_timer_* (SPLA src/timing/timing.cpp:49-86, exported by SPLA builds with
// SPLA_TIMING) and the spla_amd_timer_* extensions, over the process-wide
// timer of runtime/timer.hpp.
#include <fstream>
#include <iostream>
#include <string>

#include "api/internal_access.hpp"
#include "runtime/timer.hpp"
#include "spla_amd/spla_amd.h"

namespace {

// SPLA passes (length, pointer) pairs, as Fortran strings arrive.
bool name_of(int n, const char* name, std::string& out) {
  if (n < 0) return false;
  if (n > 0 && !name) return false;
  out.assign(name ? name : "", static_cast<size_t>(n));
  return true;
}

}  // namespace

extern "C" {

SPLA_EXPORT SplaError spla_timer_start(int n, const char* name) {
  std::string id;
  if (!name_of(n, name, id)) return n < 0 ? SPLA_INVALID_PARAMETER_ERROR : SPLA_INVALID_POINTER_ERROR;
  SPLA_C_GUARD(spla_amd::timer::start(id))
}

SPLA_EXPORT SplaError spla_timer_stop(int n, const char* name) {
  std::string id;
  if (!name_of(n, name, id)) return n < 0 ? SPLA_INVALID_PARAMETER_ERROR : SPLA_INVALID_POINTER_ERROR;
  SPLA_C_GUARD(spla_amd::timer::stop(id))
}

SPLA_EXPORT SplaError spla_timer_export_json(int n, const char* name) {
  std::string file;
  if (!name_of(n, name, file)) return n < 0 ? SPLA_INVALID_PARAMETER_ERROR : SPLA_INVALID_POINTER_ERROR;
  try {
    std::ofstream out(file);
    if (!out) return SPLA_INVALID_PARAMETER_ERROR;
    out << spla_amd::timer::json();
    if (!out) return SPLA_UNKNOWN_ERROR;
  } catch (...) {
    return SPLA_UNKNOWN_ERROR;
  }
  return SPLA_SUCCESS;
}

SPLA_EXPORT SplaError spla_timer_print() {
  SPLA_C_GUARD(std::cout << spla_amd::timer::table() << std::flush)
}

SPLA_EXPORT SplaError spla_amd_timer_set_library_scopes(int enable) {
  SPLA_C_GUARD(spla_amd::timer::set_library_scopes(enable != 0))
}

SPLA_EXPORT SplaError spla_amd_timer_clear() { SPLA_C_GUARD(spla_amd::timer::clear()) }

}  // extern "C"
