/*
 * Exception hierarchy of the C++ API. Every class maps to one SplaError code;
 * the C entry points catch GenericError and return error_code(), anything
 * else becomes SPLA_UNKNOWN_ERROR. Class names, bases and codes follow
 * SPLA 1.6.1 include/spla/exceptions.hpp:41-172.
 */
#ifndef SPLA_EXCEPTIONS_HPP
#define SPLA_EXCEPTIONS_HPP

#include <stdexcept>

#include "spla/config.h"
#include "spla/errors.h"

namespace spla {

class SPLA_EXPORT GenericError : public std::exception {
public:
  const char* what() const noexcept override { return "SPLA: Generic error"; }
  virtual SplaError error_code() const noexcept { return SplaError::SPLA_UNKNOWN_ERROR; }
};

#define SPLA_DEFINE_ERROR(NAME, BASE, TEXT, CODE)                             \
  class SPLA_EXPORT NAME : public BASE {                                      \
  public:                                                                     \
    const char* what() const noexcept override { return TEXT; }               \
    SplaError error_code() const noexcept override { return SplaError::CODE; } \
  };

SPLA_DEFINE_ERROR(InternalError, GenericError, "SPLA: Internal error", SPLA_INTERNAL_ERROR)
SPLA_DEFINE_ERROR(InvalidParameterError, GenericError, "SPLA: Invalid parameter error",
                  SPLA_INVALID_PARAMETER_ERROR)
SPLA_DEFINE_ERROR(InvalidPointerError, GenericError, "SPLA: Invalid pointer error",
                  SPLA_INVALID_POINTER_ERROR)
SPLA_DEFINE_ERROR(InvalidAllocatorFunctionError, GenericError,
                  "SPLA: Invalid allocator function error", SPLA_INVALID_ALLOCATOR_FUNCTION)
SPLA_DEFINE_ERROR(MPIError, GenericError, "SPLA: MPI error", SPLA_MPI_ERROR)
SPLA_DEFINE_ERROR(MPIAllocError, MPIError, "SPLA: MPI memory allocation error",
                  SPLA_MPI_ALLOCATION_ERROR)
SPLA_DEFINE_ERROR(MPIThreadSupportError, MPIError,
                  "SPLA: MPI multi-threading support not sufficient",
                  SPLA_MPI_THREAD_SUPPORT_ERROR)
SPLA_DEFINE_ERROR(GPUError, GenericError, "SPLA: GPU error", SPLA_GPU_ERROR)
SPLA_DEFINE_ERROR(GPUSupportError, GPUError, "SPLA: Not compiled with GPU support",
                  SPLA_GPU_SUPPORT_ERROR)
SPLA_DEFINE_ERROR(GPUAllocationError, GPUError, "SPLA: GPU allocation error",
                  SPLA_GPU_ALLOCATION_ERROR)
SPLA_DEFINE_ERROR(GPULaunchError, GPUError, "SPLA: GPU launch error", SPLA_GPU_LAUNCH_ERROR)
SPLA_DEFINE_ERROR(GPUNoDeviceError, GPUError, "SPLA: GPU no device error",
                  SPLA_GPU_NO_DEVICE_ERROR)
SPLA_DEFINE_ERROR(GPUInvalidValueError, GPUError, "SPLA: GPU invalid value error",
                  SPLA_GPU_INVALID_VALUE_ERROR)
SPLA_DEFINE_ERROR(GPUInvalidDevicePointerError, GPUError,
                  "SPLA: GPU invalid device pointer error", SPLA_GPU_INVALID_DEVICE_POINTER_ERROR)
SPLA_DEFINE_ERROR(GPUBlasError, GPUError, "SPLA: GPU BLAS error", SPLA_GPU_BLAS_ERROR)

#undef SPLA_DEFINE_ERROR

}  // namespace spla

#endif
