// Internal state of spla::Context: tunables with SPLA's defaults and checks
// (src/spla/context_internal.hpp:58-140), memory resources, HIP streams and
// the tile-plan cache. Streams are created lazily on first GPU use and reused
// across calls. A context is not thread safe (as in SPLA); the memory
// resources are.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "runtime/allocators.hpp"
#include "runtime/phase_timer.hpp"
#include "runtime/split_report.hpp"
#include "runtime/tuner.hpp"
#include "spla/types.h"

namespace spla {

class ContextInternal {
public:
  explicit ContextInternal(SplaProcessingUnit pu);
  ~ContextInternal();

  ContextInternal(const ContextInternal&) = delete;
  ContextInternal& operator=(const ContextInternal&) = delete;

  SplaProcessingUnit processing_unit() const { return pu_; }
  int num_tiles() const { return numTiles_; }
  int tile_size_host() const { return tileSizeHost_; }
  int tile_size_gpu() const { return tileSizeGPU_; }
  int op_threshold_gpu() const { return opThresholdGPU_; }
  int gpu_device_id() const { return deviceId_; }

  void set_num_tiles(int v);
  void set_tile_size_host(int v);
  void set_tile_size_gpu(int v);
  void set_op_threshold_gpu(int v);

  const std::shared_ptr<spla_amd::MemoryResource>& host_memory() const { return host_; }
  const std::shared_ptr<spla_amd::MemoryResource>& pinned_memory() const { return pinned_; }
  const std::shared_ptr<spla_amd::MemoryResource>& gpu_memory() const { return gpu_; }
  void set_host_memory(std::shared_ptr<spla_amd::MemoryResource> r) { host_ = std::move(r); }
  void set_pinned_memory(std::shared_ptr<spla_amd::MemoryResource> r) { pinned_ = std::move(r); }
  void set_gpu_memory(std::shared_ptr<spla_amd::MemoryResource> r) { gpu_ = std::move(r); }

  // Device the engine runs on: the device current at creation for a GPU
  // context, the device current at first use for a host context.
  int device();

  // Device resources; created on first use (the caller makes device() current).
  hipStream_t compute_stream();
  // Caller-provided stream for the kernels (nullptr: the context's own).
  void set_user_stream(hipStream_t s) { userStream_ = s; }

  // Target bytes of one streamed k-chunk of host-resident operands.
  std::size_t stream_chunk_bytes() const { return streamChunkBytes_; }
  void set_stream_chunk_bytes(std::size_t b) { streamChunkBytes_ = b; }
  hipStream_t copy_stream();
  // Extra streams for independent products of one batch (created on first use).
  hipStream_t batch_stream(int i);
  static constexpr int kBatchStreams = 3;
  int num_cus();

  spla_amd::Tuner& tuner() { return tuner_; }
  // FP64 matrix rate a tile of the batched GEMM sustains (fraction of the peak)
  // per tile configuration: the model's value until a measured batch plan
  // corrects it.
  double batch_efficiency(int cfg) const;
  void set_batch_efficiency(int cfg, double e);
  // dispatch of the batched GEMM launch: persistent (static schedule) or grouped
  bool batch_persistent() const { return batchPersistent_; }
  void set_batch_persistent(bool p) { batchPersistent_ = p; }

  // Arithmetic of FP64 / complex-FP64 products (include/spla_amd/spla_amd.h).
  int gemm_mode() const { return gemmMode_; }
  void set_gemm_mode(int m) { gemmMode_ = m; }
  double split_tolerance() const { return splitTolerance_; }
  void set_split_tolerance(double t) { splitTolerance_ = t; }
  int split_max_moduli() const { return splitMaxModuli_; }
  void set_split_max_moduli(int v) { splitMaxModuli_ = v; }
  // Largest device workspace (residue planes and INT8 GEMM output) of one split
  // product: the caller's value, or a quarter of the device memory (0 = default).
  std::size_t split_workspace_limit();
  void set_split_workspace_limit(std::size_t b) { splitWorkspaceLimit_ = b; }
  spla_amd::SplitReport& last_split() { return lastSplit_; }
  spla_amd::PhaseTimer& timer() { return timer_; }

private:
  void ensure_device_resources();

  SplaProcessingUnit pu_;
  int numTiles_ = 4;
  int tileSizeHost_;
  int tileSizeGPU_ = 2048;
  int opThresholdGPU_ = 2000000;
  int deviceId_ = 0;
  bool deviceKnown_ = false;

  std::shared_ptr<spla_amd::MemoryResource> host_;
  std::shared_ptr<spla_amd::MemoryResource> pinned_;
  std::shared_ptr<spla_amd::MemoryResource> gpu_;

  hipStream_t compute_ = nullptr;
  hipStream_t userStream_ = nullptr;
  std::size_t streamChunkBytes_ = std::size_t(256) << 20;
  hipStream_t copy_ = nullptr;
  hipStream_t batch_[kBatchStreams] = {nullptr, nullptr, nullptr};
  int numCUs_ = 0;
  std::size_t deviceMemory_ = 0;

  spla_amd::Tuner tuner_;
  double batchEfficiency_[2] = {0.75, 0.55};  // TileL, TileM
  bool batchPersistent_ = false;
  int gemmMode_ = 0;
  double splitTolerance_ = 1e-14;
  int splitMaxModuli_ = 20;
  std::size_t splitWorkspaceLimit_ = 0;
  spla_amd::SplitReport lastSplit_;
  spla_amd::PhaseTimer timer_;
};

}  // namespace spla
