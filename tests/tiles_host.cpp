// Host harness for tests/test_tile_order.py: the launch orders of orctm_amd/csrc/nh_tiles.h, compiled
// with g++, on tile lists described by a system index per tile and the systems' depths.
#include <cstdint>
#include <vector>

#include "../orctm_amd/csrc/nh_tiles.h"

namespace {
struct T {
  int sys, bx, by, local;
};
std::vector<T> tiles(const int32_t* sys, int n) {
  std::vector<T> v(n);
  for (int k = 0; k < n; ++k) v[k] = T{sys[k], 0, 0, k};
  return v;
}
}  // namespace

// out[b] = the natural index of the tile launched as block b, or -1 for an empty tile; returns the
// list's length (at most out_cap).
extern "C" int balanced_order(const int32_t* sys, int n, const int32_t* depth, int min_tiles, int32_t* out,
                              int out_cap) {
  auto v = onh::balanced_order(tiles(sys, n), [&](const T& t) { return depth[t.sys]; }, min_tiles, T{-1, 0, 0, -1});
  if ((int)v.size() > out_cap) return -1;
  for (size_t b = 0; b < v.size(); ++b) out[b] = v[b].local;
  return (int)v.size();
}

extern "C" int xcd_order(const int32_t* sys, int n, int32_t* out) {
  auto v = onh::xcd_order(tiles(sys, n));
  for (size_t b = 0; b < v.size(); ++b) out[b] = v[b].local;
  return (int)v.size();
}
