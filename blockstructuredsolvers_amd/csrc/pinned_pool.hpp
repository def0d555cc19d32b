// Process-wide pool of small pinned host buffers (status words and backward errors
// that the device copies back after a factorisation or a refinement check).
//
// Plans are short-lived where requests are served (one group plan per admitted group),
// and hipHostMalloc / hipHostFree per plan cost milliseconds: freeing pinned memory
// waits for the device.  Plans take their buffers from this pool and give them back
// when destroyed; the pool keeps what was in use at the peak until the process ends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace bss {

struct PinnedPool {
  std::mutex mu;
  std::map<size_t, std::vector<void*>> free_by_class;  // size class -> idle buffers
  std::unordered_map<void*, size_t> class_of;           // every buffer the pool made
};

// never destroyed: a pinned buffer must not be released while the runtime shuts down
inline PinnedPool& pinned_pool() {
  static PinnedPool* pool = new PinnedPool;
  return *pool;
}

// size classes: powers of two from 256 bytes
inline size_t pinned_class(size_t bytes) {
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}

// A pinned buffer of at least `bytes` bytes (contents undefined), or nullptr.
inline void* pinned_take(size_t bytes) {
  const size_t c = pinned_class(bytes);
  PinnedPool& P = pinned_pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.free_by_class.find(c);
  if (it != P.free_by_class.end() && !it->second.empty()) {
    void* q = it->second.back();
    it->second.pop_back();
    return q;
  }
  void* q = nullptr;
  if (hipHostMalloc(&q, c) != hipSuccess) return nullptr;
  P.class_of[q] = c;
  return q;
}

// Returns a buffer from pinned_take to the pool; the caller must have no copy into or
// out of it in flight.
inline void pinned_give(void* q) {
  if (!q) return;
  PinnedPool& P = pinned_pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.class_of.find(q);
  if (it != P.class_of.end()) P.free_by_class[it->second].push_back(q);
}

}  // namespace bss
