// nh_pool.cpp -- the library's device-memory pool (nh_pool.h).
//
// Per device, slabs from hipMalloc are carved first-fit in address order; freed blocks coalesce with
// their free neighbours inside their slab (never across slabs: two hipMalloc ranges that happen to be
// adjacent in the address space are still two allocations to the runtime's copy engines). Slabs are
// kept until the process ends.
#include "nh_pool.h"

#include <algorithm>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace onh {
namespace {

constexpr size_t kLarge = size_t(1) << 20;       // blocks of 1 MiB or more are large
constexpr size_t kPage = size_t(2) << 20;        // large blocks: 2 MiB aligned, 2 MiB granules
constexpr size_t kSmallAlign = 256;              // small blocks: 256-byte granules
constexpr size_t kSmallSlab = size_t(64) << 20;  // one slab holds many small blocks
constexpr size_t kLargeSlab = size_t(1) << 30;   // least size of a slab of large blocks

constexpr size_t kHostSlab = size_t(4) << 20;    // one slab of pinned host blocks
constexpr int kHost = -1;                        // pool id of the pinned host blocks

struct Slab {
  void* raw = nullptr;  // what hipMalloc returned (freed with the process)
  char* base = nullptr; // first usable byte (2 MiB aligned for large slabs)
  size_t size = 0;
  bool small = false;
  std::map<size_t, size_t> idle;  // offset -> length, coalesced
};

struct DevicePool {
  std::vector<std::unique_ptr<Slab>> slabs;
  std::unordered_map<char*, std::pair<Slab*, size_t>> live;  // block -> (slab, length)
  size_t held = 0, used = 0;
};

std::mutex g_mu;
std::map<int, DevicePool> g_pools;

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

bool take(Slab& s, size_t len, size_t* off) {
  for (auto it = s.idle.begin(); it != s.idle.end(); ++it) {
    if (it->second < len) continue;
    *off = it->first;
    const size_t rest = it->second - len, at = it->first + len;
    s.idle.erase(it);
    if (rest) s.idle[at] = rest;
    return true;
  }
  return false;
}

void give(Slab& s, size_t off, size_t len) {
  auto next = s.idle.lower_bound(off);
  if (next != s.idle.end() && off + len == next->first) {
    len += next->second;
    next = s.idle.erase(next);
  }
  if (next != s.idle.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      prev->second += len;
      return;
    }
  }
  s.idle[off] = len;
}

}  // namespace

hipError_t pool_alloc(int device, size_t bytes, void** p) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  const bool host = device == kHost;
  if (host && bytes > kHostSlab) return hipErrorInvalidValue;
  const bool small = host || bytes < kLarge;
  const size_t len = small ? round_up(bytes, kSmallAlign) : round_up(bytes, kPage);
  std::lock_guard<std::mutex> lk(g_mu);
  DevicePool& d = g_pools[device];
  for (auto& sp : d.slabs) {
    Slab& s = *sp;
    size_t off = 0;
    if (s.small != small || !take(s, len, &off)) continue;
    char* b = s.base + off;
    d.live[b] = {&s, len};
    d.used += len;
    *p = b;
    return hipSuccess;
  }
  // grow: a new slab (large slabs get a spare 2 MiB so that their base can be aligned), on `device`
  // whatever device the calling thread has current
  int cur = -1;
  if (!host && hipGetDevice(&cur) == hipSuccess && cur != device) {
    const hipError_t es = hipSetDevice(device);
    if (es != hipSuccess) return es;
  }
  size_t want = host ? kHostSlab : small ? kSmallSlab : std::max(len, kLargeSlab);
  void* raw = nullptr;
  hipError_t e = host ? hipHostMalloc(&raw, want, hipHostMallocDefault) : hipMalloc(&raw, small ? want : want + kPage);
  if (e != hipSuccess && !small && want > len) {
    (void)hipGetLastError();
    want = len;
    e = hipMalloc(&raw, want + kPage);
  }
  if (cur >= 0 && cur != device) hipSetDevice(cur);
  if (e != hipSuccess) return e;
  auto s = std::make_unique<Slab>();
  s->raw = raw;
  s->small = small;
  s->size = want;
  const uintptr_t a = reinterpret_cast<uintptr_t>(raw);
  s->base = small ? static_cast<char*>(raw) : reinterpret_cast<char*>(round_up(a, kPage));
  s->idle[0] = want;
  d.held += small ? want : want + kPage;
  size_t off = 0;
  take(*s, len, &off);
  char* b = s->base + off;
  d.live[b] = {s.get(), len};
  d.used += len;
  d.slabs.push_back(std::move(s));
  *p = b;
  return hipSuccess;
}

void pool_free(int device, void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_mu);
  auto pit = g_pools.find(device);
  if (pit == g_pools.end()) return;
  DevicePool& d = pit->second;
  auto it = d.live.find(static_cast<char*>(p));
  if (it == d.live.end()) return;
  Slab& s = *it->second.first;
  const size_t len = it->second.second;
  give(s, static_cast<size_t>(static_cast<char*>(p) - s.base), len);
  d.used -= len;
  d.live.erase(it);
}

hipError_t pool_alloc_host(size_t bytes, void** p) { return pool_alloc(kHost, bytes, p); }

void pool_free_host(void* p) { pool_free(kHost, p); }

size_t pool_idle_bytes(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_pools.find(device);
  if (it == g_pools.end()) return 0;
  size_t n = 0;
  for (auto& s : it->second.slabs)
    for (auto& kv : s->idle) n += kv.second;
  return n;
}

size_t pool_held_bytes(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_pools.find(device);
  return it == g_pools.end() ? 0 : it->second.held;
}

}  // namespace onh
