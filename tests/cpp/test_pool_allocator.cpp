// Checks of the caching pool (spla_amd/csrc/runtime/pool_allocator.hpp), after SPLA's
// tests/test_pool_allocator.cpp:17-109: serial, interleaved and 100k random
// allocate / deallocate sequences against an instrumented backend that detects
// leaks, double frees, foreign frees and blocks smaller than requested.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "runtime/pool_allocator.hpp"

namespace {
std::map<void*, std::size_t> g_live;
int g_errors = 0;
long g_backend_allocs = 0;

void* backend_alloc(std::size_t n) {
  void* p = std::malloc(n);
  g_live[p] = n;
  ++g_backend_allocs;
  return p;
}
void backend_free(void* p) {
  auto it = g_live.find(p);
  if (it == g_live.end()) {
    std::printf("double or foreign free %p\n", p);
    ++g_errors;
    return;
  }
  g_live.erase(it);
  std::free(p);
}
void check(bool c, const char* what) {
  if (!c) {
    std::printf("FAILED: %s\n", what);
    ++g_errors;
  }
}
std::size_t held_bytes() {
  std::size_t s = 0;
  for (auto& kv : g_live) s += kv.second;
  return s;
}
}  // namespace

int main() {
  {
    spla_amd::PoolAllocator pool(backend_alloc, backend_free);
    check(pool.allocate(0) == nullptr, "allocate(0) is null");
    // serial
    for (int i = 1; i <= 100; ++i) {
      void* p = pool.allocate(i * 10);
      check(p != nullptr, "serial allocation");
      check(g_live.count(p) && g_live[p] >= std::size_t(i * 10), "block >= request");
      pool.deallocate(p);
    }
    check(pool.size() == held_bytes(), "size() == bytes held (serial)");
    // reuse: best fit hands back a cached block without a backend call
    long before = g_backend_allocs;
    void* a = pool.allocate(500);
    check(g_backend_allocs == before, "cached block reused");
    pool.deallocate(a);
    // interleaved
    std::vector<void*> ps;
    for (int i = 0; i < 50; ++i) ps.push_back(pool.allocate(64 + i));
    for (int i = 0; i < 50; i += 2) pool.deallocate(ps[i]);
    for (int i = 0; i < 25; ++i) ps.push_back(pool.allocate(32 + 3 * i));
    for (std::size_t i = 1; i < 50; i += 2) pool.deallocate(ps[i]);
    for (std::size_t i = 50; i < ps.size(); ++i) pool.deallocate(ps[i]);
    check(pool.size() == held_bytes(), "size() == bytes held (interleaved)");
    pool.deallocate(reinterpret_cast<void*>(0x1234));  // unknown pointer: ignored
    // random
    std::mt19937 gen(42);
    std::uniform_int_distribution<int> coin(0, 2), sz(1, 1 << 16);
    std::vector<std::pair<void*, std::size_t>> live;
    for (int it = 0; it < 100000; ++it) {
      if (live.empty() || coin(gen) != 0) {
        std::size_t n = sz(gen);
        void* p = pool.allocate(n);
        if (!g_live.count(p) || g_live[p] < n) {
          check(false, "random: block >= request");
        }
        live.emplace_back(p, n);
      } else {
        std::uniform_int_distribution<std::size_t> pick(0, live.size() - 1);
        std::size_t i = pick(gen);
        pool.deallocate(live[i].first);
        live[i] = live.back();
        live.pop_back();
      }
      if (live.size() > 200) {
        for (auto& kv : live) pool.deallocate(kv.first);
        live.clear();
      }
    }
    // distinct live blocks never alias
    std::map<void*, int> seen;
    for (auto& kv : live) check(seen[kv.first]++ == 0, "live blocks distinct");
    for (auto& kv : live) pool.deallocate(kv.first);
    check(pool.size() == held_bytes(), "size() == bytes held (random)");
    pool.release();
    check(pool.size() == 0 && g_live.empty(), "release() frees every cached block");
    void* b = pool.allocate(100);
    pool.deallocate(b);
  }
  check(g_live.empty(), "no leak after destruction");
  // backend failure: the pool drops its cache and retries once, then throws
  {
    int fail = 0;
    spla_amd::PoolAllocator pool(
        [&](std::size_t n) -> void* { return fail-- > 0 ? nullptr : backend_alloc(n); },
        backend_free);
    void* p = pool.allocate(10);
    pool.deallocate(p);
    fail = 1;
    void* q = pool.allocate(1000);  // first attempt fails, cache released, retry succeeds
    check(q != nullptr, "retry after releasing the cache");
    fail = 2;
    bool threw = false;
    try {
      pool.allocate(5000);
    } catch (const std::bad_alloc&) {
      threw = true;
    }
    check(threw, "bad_alloc when the backend keeps failing");
    pool.deallocate(q);
  }
  check(g_live.empty(), "no leak after failure handling");
  std::printf("pool allocator checks: %s (%d errors)\n", g_errors ? "FAILED" : "OK", g_errors);
  return g_errors ? 1 : 0;
}
