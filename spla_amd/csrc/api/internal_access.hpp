// Access to the private state of the public C++ classes from the library.
#pragma once

#include <memory>

#include "runtime/context_internal.hpp"
#include "runtime/matrix_distribution_internal.hpp"
#include "spla/context.hpp"
#include "spla/exceptions.hpp"
#include "spla/matrix_distribution.hpp"

namespace spla {

struct InternalAccess {
  static ContextInternal& ctx(Context& c) { return *c.ctxInternal_; }
  static std::shared_ptr<ContextInternal> ctx_ptr(Context& c) { return c.ctxInternal_; }
  static MatrixDistributionInternal& dist(MatrixDistribution& d) { return *d.descInternal_; }
  static MatrixDistribution make_dist(MatrixDistributionInternal&& d) {
    return MatrixDistribution(std::make_shared<MatrixDistributionInternal>(std::move(d)));
  }
};

}  // namespace spla

// Body of a C entry point: exceptions become error codes
// (SPLA's C wrappers, e.g. src/spla/pgemm_ssbtr.cpp:129-138).
#define SPLA_C_GUARD(...)                          \
  try {                                            \
    __VA_ARGS__;                                   \
  } catch (const spla::GenericError& e) {          \
    return e.error_code();                         \
  } catch (...) {                                  \
    return SplaError::SPLA_UNKNOWN_ERROR;          \
  }                                                \
  return SplaError::SPLA_SUCCESS;
