// Device memory cache behind DevBuf.
//
// Plans own many buffers (point copies, membership lists, per-pair slots)
// sized by the workload. Coupling loops and benchmarks build and destroy
// plans of the same sizes again and again; hipFree of hundreds of MB costs
// milliseconds (the runtime synchronises and unmaps), and so does mapping
// fresh memory. Freed blocks are therefore kept per device and handed out
// again for requests of the same rounded size.
//
// A freed block may still be read or written by work in flight (DevBuf is
// released when its owner goes out of scope, not when its stream is done).
// It is parked as "pending"; the first request that would reuse a pending
// block synchronises the device once, after which every pending block is
// reusable. A failed hipMalloc releases the whole cache and retries.
// PACMAN_AMD_NO_MEMORY_CACHE=1 in the environment turns the cache off; a
// released block then goes to hipFree, which waits for the device.
//
// Callers rely on these two rules: the setup code (scan_device, build_grid,
// make_centers, patch_sources, patch_targets, build_adaptive) releases its
// temporaries while the kernels that read them may still run, with no
// stream synchronisation of its own. A stream-ordered allocator
// (hipMallocAsync / hipFreeAsync) in place of this cache would keep that
// safe only on one stream; any other change has to revisit those call
// sites.

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "rbf_internal.hpp"

namespace pacman_amd {

namespace {

struct DeviceCache {
  std::multimap<size_t, void *> ready;    // size -> block
  std::multimap<size_t, void *> pending;  // freed, maybe still in use
  size_t bytes = 0;                       // held by the cache
};

std::mutex g_mu;
std::map<int, DeviceCache> g_cache;

bool cache_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("PACMAN_AMD_NO_MEMORY_CACHE");
    return !(e && e[0] == '1');
  }();
  return on;
}

// Requests are rounded so that nearly equal sizes share blocks.
size_t round_size(size_t n) {
  const size_t g = n < (1u << 20) ? 512 : (size_t)2 << 20;
  return (n + g - 1) / g * g;
}

void release_all_locked(int dev, DeviceCache &c) {
  (void)dev;
  if (!c.pending.empty())
    (void)hipDeviceSynchronize();
  for (auto &kv : c.ready)
    (void)hipFree(kv.second);
  for (auto &kv : c.pending)
    (void)hipFree(kv.second);
  c.ready.clear();
  c.pending.clear();
  c.bytes = 0;
}

} // namespace

void *dev_malloc(size_t bytes) {
  if (bytes == 0)
    return nullptr;
  if (!cache_enabled()) {
    void *p = nullptr;
    PACMAN_HIP_CHECK(hipMalloc(&p, bytes));
    return p;
  }
  const size_t n = round_size(bytes);
  int dev = 0;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceCache &c = g_cache[dev];
  auto take = [&](std::multimap<size_t, void *> &m) -> void * {
    auto it = m.find(n);
    if (it == m.end())
      return nullptr;
    void *p = it->second;
    m.erase(it);
    c.bytes -= n;
    return p;
  };
  if (void *p = take(c.ready))
    return p;
  if (c.pending.count(n)) {
    // everything freed so far becomes safe to reuse
    PACMAN_HIP_CHECK(hipDeviceSynchronize());
    c.ready.insert(c.pending.begin(), c.pending.end());
    c.pending.clear();
    if (void *p = take(c.ready))
      return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, n) != hipSuccess) {
    (void)hipGetLastError();
    release_all_locked(dev, c);
    PACMAN_HIP_CHECK(hipMalloc(&p, n));
  }
  return p;
}

void dev_free(void *p, size_t bytes) {
  if (p == nullptr)
    return;
  if (!cache_enabled()) {
    (void)hipFree(p);
    return;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return;
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceCache &c = g_cache[dev];
  const size_t n = round_size(bytes);
  c.pending.emplace(n, p);
  c.bytes += n;
}

void dev_release_cache() {
  std::lock_guard<std::mutex> lk(g_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_cache) {
    (void)hipSetDevice(kv.first);
    release_all_locked(kv.first, kv.second);
  }
  (void)hipSetDevice(cur);
}

size_t dev_cached_bytes() {
  std::lock_guard<std::mutex> lk(g_mu);
  size_t s = 0;
  for (auto &kv : g_cache)
    s += kv.second.bytes;
  return s;
}

} // namespace pacman_amd
