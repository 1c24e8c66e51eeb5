// Best-fit caching pool over an arbitrary allocate / deallocate pair.
//
// Observable behaviour follows SPLA's PoolAllocator
// (src/memory/pool_allocator.hpp:80-124):
//   * allocate(0) returns nullptr;
//   * a request is served by the smallest free block whose size is >= the
//     request (the block is handed out whole);
//   * when no free block is large enough, the largest free block is returned
//     to the backend and a block of exactly the requested size is allocated;
//   * size() is the number of bytes currently held from the backend
//     (handed out + cached).
// Additions for large device pools: when the backend allocation fails, every
// cached block is released and the allocation is retried once; release()
// empties the cache. The pool is thread safe. Header only, so that the host
// logic can be built and tested without a GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
#include <utility>

namespace spla_amd {

class PoolAllocator {
public:
  using AllocFn = std::function<void*(std::size_t)>;
  using FreeFn = std::function<void(void*)>;

  PoolAllocator(AllocFn allocateFunc, FreeFn deallocateFunc)
      : alloc_(std::move(allocateFunc)), free_(std::move(deallocateFunc)) {
    if (!alloc_ || !free_) throw std::invalid_argument("pool allocator needs both functions");
  }
  PoolAllocator(const PoolAllocator&) = delete;
  PoolAllocator& operator=(const PoolAllocator&) = delete;

  ~PoolAllocator() {
    for (auto& kv : used_) free_(kv.first);
    for (auto& kv : cached_) free_(kv.second);
  }

  void* allocate(std::size_t size) {
    if (size == 0) return nullptr;
    std::lock_guard<std::mutex> guard(lock_);
    auto fit = cached_.lower_bound(size);
    if (fit != cached_.end()) {
      void* ptr = fit->second;
      used_.emplace(ptr, fit->first);
      cached_.erase(fit);
      return ptr;
    }
    if (!cached_.empty()) {
      auto largest = std::prev(cached_.end());
      free_(largest->second);
      held_ -= largest->first;
      cached_.erase(largest);
    }
    void* ptr = alloc_(size);
    if (!ptr) {
      release_cached_locked();
      ptr = alloc_(size);
    }
    if (!ptr) throw std::bad_alloc();
    held_ += size;
    used_.emplace(ptr, size);
    return ptr;
  }

  // Returns a block to the cache. Unknown pointers are ignored (never throws).
  void deallocate(void* ptr) noexcept {
    if (!ptr) return;
    std::lock_guard<std::mutex> guard(lock_);
    auto it = used_.find(ptr);
    if (it == used_.end()) return;
    cached_.emplace(it->second, it->first);
    used_.erase(it);
  }

  // Frees every cached (not handed out) block.
  void release() {
    std::lock_guard<std::mutex> guard(lock_);
    release_cached_locked();
  }

  std::uint_least64_t size() {
    std::lock_guard<std::mutex> guard(lock_);
    return held_;
  }

  std::size_t num_cached() {
    std::lock_guard<std::mutex> guard(lock_);
    return cached_.size();
  }

  std::size_t num_used() {
    std::lock_guard<std::mutex> guard(lock_);
    return used_.size();
  }

private:
  void release_cached_locked() {
    for (auto& kv : cached_) {
      free_(kv.second);
      held_ -= kv.first;
    }
    cached_.clear();
  }

  AllocFn alloc_;
  FreeFn free_;
  std::multimap<std::size_t, void*> cached_;
  std::unordered_map<void*, std::size_t> used_;
  std::mutex lock_;
  std::uint_least64_t held_ = 0;
};

}  // namespace spla_amd
