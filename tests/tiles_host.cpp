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

// The paired column update's blocks for systems of tx[q] x ty[q] tiles (natural order): out holds
// (system, bx, by, vertical) per block; returns the number of blocks (at most out_cap).
extern "C" int pair_tiles(const int32_t* tx, const int32_t* ty, int nsys, int32_t* out, int out_cap) {
  std::vector<T> nat;
  for (int q = 0; q < nsys; ++q)
    for (int by = 0; by < ty[q]; ++by)
      for (int bx = 0; bx < tx[q]; ++bx) nat.push_back(T{q, bx, by, (int)nat.size()});
  auto v = onh::pair_tiles(nat);
  if ((int)v.size() > out_cap) return -1;
  for (size_t b = 0; b < v.size(); ++b) {
    out[4 * b] = v[b].sys;
    out[4 * b + 1] = v[b].bx;
    out[4 * b + 2] = v[b].by;
    out[4 * b + 3] = v[b].local;
  }
  return (int)v.size();
}
