// Device memory of the library: every DevBuf allocation and release goes
// through device_alloc / device_free.
//
// By default a released block goes back to the driver at once (hipFree).
// With a cache limit set, released blocks are kept per device, up to that
// many bytes, and a later allocation of the library reuses one of the same
// size (at most 25% larger).  A loop that builds a grid on the device, drops
// the previous one and builds again then stops waiting on the driver, whose
// hipMalloc of recently freed memory can take seconds at hundreds of GB.
//
// A block enters the cache only after a device-wide synchronization (the
// wait hipFree itself does), so no kernel still in flight on any stream can
// reach a block that another object takes next.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "casl_reinit.h"
#include "reinit_device.hpp"

namespace casl {
namespace {

struct BlockCache {
  std::mutex mu;
  size_t limit = 0;                                     // bytes per device; 0: no cache
  std::map<int, size_t> held;                           // device -> bytes in the cache
  std::map<int, std::multimap<size_t, void *>> blocks;  // device -> size -> block
};

// never destroyed: blocks released during static destruction at exit still
// find it (the driver reclaims whatever it holds then)
BlockCache &block_cache() {
  static BlockCache *c = new BlockCache;
  return *c;
}

// allocation sizes: 256-byte granules, 2 MiB ones from 1 MiB up (so that the
// slightly different sizes of consecutive grids map to the same blocks)
size_t granted_size(size_t bytes) {
  const size_t g = bytes >= (size_t(1) << 20) ? (size_t(2) << 20) : size_t(256);
  return (bytes + g - 1) / g * g;
}

// frees every cached block of one device; the caller holds the lock
void trim_locked(BlockCache &c, int device) {
  auto it = c.blocks.find(device);
  if (it == c.blocks.end()) return;
  for (auto &kv : it->second) (void)hipFree(kv.second);
  it->second.clear();
  c.held[device] = 0;
}

}  // namespace

void *device_alloc(size_t bytes, size_t *granted, int *device) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  const size_t want = granted_size(bytes);
  BlockCache &c = block_cache();
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto bi = c.blocks.find(dev);
    if (bi != c.blocks.end()) {
      auto it = bi->second.lower_bound(want);
      if (it != bi->second.end() && it->first <= want + want / 4) {
        void *p = it->second;
        *granted = it->first;
        *device = dev;
        c.held[dev] -= it->first;
        bi->second.erase(it);
        return p;
      }
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();  // the failure is handled here, not left for a later launch check
    std::lock_guard<std::mutex> g(c.mu);
    trim_locked(c, dev);
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    throw hip_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  *granted = want;
  *device = dev;
  return p;
}

void device_free(void *p, size_t bytes, int device) {
  if (!p) return;
  BlockCache &c = block_cache();
  std::lock_guard<std::mutex> g(c.mu);
  if (c.limit == 0 || c.held[device] + bytes > c.limit) {
    (void)hipFree(p);
    return;
  }
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || (cur != device && hipSetDevice(device) != hipSuccess) ||
      hipDeviceSynchronize() != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(p);
  } else {
    c.blocks[device].emplace(bytes, p);
    c.held[device] += bytes;
  }
  if (cur != device) (void)hipSetDevice(cur);
}

}  // namespace casl

extern "C" {

int casl_device_cache_set_limit(size_t bytes_per_device) {
  return casl::guarded([&]() {
    casl::BlockCache &c = casl::block_cache();
    std::lock_guard<std::mutex> g(c.mu);
    c.limit = bytes_per_device;
    for (auto &kv : c.held)
      if (kv.second > c.limit) casl::trim_locked(c, kv.first);
  });
}

int casl_device_cache_trim(void) {
  return casl::guarded([&]() {
    casl::BlockCache &c = casl::block_cache();
    std::lock_guard<std::mutex> g(c.mu);
    for (auto &kv : c.held) casl::trim_locked(c, kv.first);
  });
}

int casl_device_cache_held(int device, size_t *bytes) {
  return casl::guarded([&]() {
    if (!bytes) throw std::invalid_argument("null output");
    casl::BlockCache &c = casl::block_cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.held.find(device);
    *bytes = it == c.held.end() ? 0 : it->second;
  });
}

}  // extern "C"
