#include "runtime/context_internal.hpp"

#include <cstdlib>

#include "runtime/hip_check.hpp"
#include "spla/exceptions.hpp"

namespace spla {

ContextInternal::ContextInternal(SplaProcessingUnit pu)
    : pu_(pu), tileSizeHost_(pu == SPLA_PU_HOST ? 500 : 1500) {
  if (pu != SPLA_PU_HOST && pu != SPLA_PU_GPU) throw InvalidParameterError();
  host_ = std::make_shared<spla_amd::PooledResource>(
      [](std::size_t s) -> void* { return std::malloc(s); }, [](void* p) { std::free(p); },
      [] { throw InternalError(); });
  pinned_ = std::make_shared<spla_amd::PooledResource>(
      [](std::size_t s) -> void* {
        void* p = nullptr;
        if (hipHostMalloc(&p, s, hipHostMallocDefault) != hipSuccess) {
          (void)hipGetLastError();
          return nullptr;
        }
        return p;
      },
      [](void* p) { (void)hipHostFree(p); }, [] { throw GPUAllocationError(); });
  gpu_ = std::make_shared<spla_amd::PooledResource>(
      [](std::size_t s) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, s) != hipSuccess) {
          (void)hipGetLastError();
          return nullptr;
        }
        return p;
      },
      [](void* p) { (void)hipFree(p); }, [] { throw GPUAllocationError(); });
  if (pu == SPLA_PU_GPU) {
    // SPLA captures the current device at creation (context_internal.hpp:66-69).
    spla_amd::hip_check(hipGetDevice(&deviceId_));
    deviceKnown_ = true;
  }
}

ContextInternal::~ContextInternal() {
  // Pools release their blocks in their destructors; the streams go first.
  if (compute_) (void)hipStreamDestroy(compute_);
  if (copy_) (void)hipStreamDestroy(copy_);
  for (hipStream_t b : batch_)
    if (b) (void)hipStreamDestroy(b);
}

void ContextInternal::set_num_tiles(int v) {
  if (v < 1) throw InvalidParameterError();
  numTiles_ = v;
}
void ContextInternal::set_tile_size_host(int v) {
  if (v < 1) throw InvalidParameterError();
  tileSizeHost_ = v;
}
void ContextInternal::set_tile_size_gpu(int v) {
  if (v < 1) throw InvalidParameterError();
  tileSizeGPU_ = v;
}
void ContextInternal::set_op_threshold_gpu(int v) {
  if (v < 0) throw InvalidParameterError();
  opThresholdGPU_ = v;
}

int ContextInternal::device() {
  if (!deviceKnown_) {
    // A host-PU context is served by the GPU engine on the current device.
    spla_amd::hip_check(hipGetDevice(&deviceId_));
    deviceKnown_ = true;
  }
  return deviceId_;
}

void ContextInternal::ensure_device_resources() {
  device();
  if (!compute_) {
    // Default priority only: streams of different priorities are avoided.
    spla_amd::hip_check(hipStreamCreateWithFlags(&compute_, hipStreamNonBlocking));
    spla_amd::hip_check(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    spla_amd::hip_check(hipGetDeviceProperties(&prop, deviceId_));
    numCUs_ = prop.multiProcessorCount;
    deviceMemory_ = prop.totalGlobalMem;
  }
}

std::size_t ContextInternal::split_workspace_limit() {
  if (splitWorkspaceLimit_) return splitWorkspaceLimit_;
  ensure_device_resources();
  return deviceMemory_ / 4;
}

hipStream_t ContextInternal::compute_stream() {
  ensure_device_resources();
  return userStream_ ? userStream_ : compute_;
}

hipStream_t ContextInternal::copy_stream() {
  ensure_device_resources();
  return copy_;
}

hipStream_t ContextInternal::batch_stream(int i) {
  ensure_device_resources();
  if (!batch_[i])  // default priority, like every stream of the engine
    spla_amd::hip_check(hipStreamCreateWithFlags(&batch_[i], hipStreamNonBlocking));
  return batch_[i];
}

double ContextInternal::batch_efficiency(int cfg) const {
  return batchEfficiency_[cfg == 1 ? 1 : 0];
}

void ContextInternal::set_batch_efficiency(int cfg, double e) {
  batchEfficiency_[cfg == 1 ? 1 : 0] = e;
}

int ContextInternal::num_cus() {
  ensure_device_resources();
  return numCUs_;
}

}  // namespace spla
