// Maps HIP runtime status codes to the exception types of the public API
// (the mapping of SPLA's src/gpu_util/gpu_runtime_api.hpp:113-123).
#pragma once

#include <hip/hip_runtime.h>

#include "spla/exceptions.hpp"

namespace spla_amd {

inline void hip_check(hipError_t status) {
  if (status == hipSuccess) return;
  (void)hipGetLastError();  // clear the sticky per-thread status
  switch (status) {
    case hipErrorMemoryAllocation:
      throw spla::GPUAllocationError();
    case hipErrorLaunchFailure:
    case hipErrorLaunchOutOfResources:
    case hipErrorInvalidConfiguration:
      throw spla::GPULaunchError();
    case hipErrorNoDevice:
      throw spla::GPUNoDeviceError();
    case hipErrorInvalidValue:
      throw spla::GPUInvalidValueError();
    case hipErrorInvalidDevicePointer:
      throw spla::GPUInvalidDevicePointerError();
    default:
      throw spla::GPUError();
  }
}

}  // namespace spla_amd
