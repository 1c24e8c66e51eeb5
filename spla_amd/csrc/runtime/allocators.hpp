// Memory resources of a context: host, pinned host and device memory.
//
// Each resource is either a caching pool (default) or a pair of user
// functions installed through spla_ctx_set_alloc_* (SPLA's AllocatorCollection
// and SimpleAllocator, src/memory/allocator_collection.hpp:52-123 and
// src/memory/simple_allocator.hpp:46-101).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_map>
#include <utility>

#include "runtime/pool_allocator.hpp"

namespace spla_amd {

class MemoryResource {
public:
  virtual ~MemoryResource() = default;
  virtual void* allocate(std::size_t size) = 0;
  virtual void deallocate(void* ptr) noexcept = 0;
  virtual std::uint_least64_t size() = 0;  // bytes held
};

class PooledResource final : public MemoryResource {
public:
  // onFailure is called (and expected to throw) when the backend cannot
  // provide the memory.
  PooledResource(PoolAllocator::AllocFn a, PoolAllocator::FreeFn f,
                 std::function<void()> onFailure = nullptr)
      : pool_(std::move(a), std::move(f)), onFailure_(std::move(onFailure)) {}
  void* allocate(std::size_t size) override {
    try {
      return pool_.allocate(size);
    } catch (const std::bad_alloc&) {
      if (onFailure_) onFailure_();
      throw;
    }
  }
  void deallocate(void* ptr) noexcept override { pool_.deallocate(ptr); }
  std::uint_least64_t size() override { return pool_.size(); }

  void release() { pool_.release(); }

private:
  PoolAllocator pool_;
  std::function<void()> onFailure_;
};

// Direct user functions; size() counts the bytes currently handed out.
class UserResource final : public MemoryResource {
public:
  UserResource(std::function<void*(std::size_t)> a, std::function<void(void*)> f)
      : alloc_(std::move(a)), free_(std::move(f)) {}
  void* allocate(std::size_t size) override {
    if (!size) return nullptr;
    std::lock_guard<std::mutex> guard(lock_);
    void* ptr = alloc_(size);
    if (!ptr) throw std::bad_alloc();
    live_.emplace(ptr, size);
    held_ += size;
    return ptr;
  }
  void deallocate(void* ptr) noexcept override {
    if (!ptr) return;
    std::lock_guard<std::mutex> guard(lock_);
    auto it = live_.find(ptr);
    if (it == live_.end()) return;
    free_(ptr);
    held_ -= it->second;
    live_.erase(it);
  }
  std::uint_least64_t size() override {
    std::lock_guard<std::mutex> guard(lock_);
    return held_;
  }

private:
  std::function<void*(std::size_t)> alloc_;
  std::function<void(void*)> free_;
  std::mutex lock_;
  std::unordered_map<void*, std::size_t> live_;
  std::uint_least64_t held_ = 0;
};

// RAII block from a resource; the resource must outlive the block.
class Block {
public:
  Block() = default;
  Block(std::shared_ptr<MemoryResource> res, std::size_t bytes) : res_(std::move(res)) {
    if (bytes) ptr_ = res_->allocate(bytes);
    bytes_ = bytes;
  }
  Block(const Block&) = delete;
  Block& operator=(const Block&) = delete;
  Block(Block&& o) noexcept { *this = std::move(o); }
  Block& operator=(Block&& o) noexcept {
    reset();
    res_ = std::move(o.res_);
    ptr_ = o.ptr_;
    bytes_ = o.bytes_;
    o.ptr_ = nullptr;
    o.bytes_ = 0;
    return *this;
  }
  ~Block() { reset(); }
  void reset() noexcept {
    if (ptr_ && res_) res_->deallocate(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
  template <typename T = void>
  T* get() const {
    return static_cast<T*>(ptr_);
  }
  std::size_t bytes() const { return bytes_; }

private:
  std::shared_ptr<MemoryResource> res_;
  void* ptr_ = nullptr;
  std::size_t bytes_ = 0;
};

}  // namespace spla_amd
