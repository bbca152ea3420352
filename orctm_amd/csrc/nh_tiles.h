// nh_tiles.h -- the order in which a launch's tiles meet the XCDs (host code, no HIP; the host harness
// tests/tiles_host.cpp compiles it with g++).
//
// Block b of a launch runs on XCD b % 8 (round-robin dispatch), and each XCD then works through its
// blocks b, b + 8, b + 16, ... on its own. Tiles that neighbour each other (halo rows and columns)
// should run on one XCD, where they share an L2; and a launch ends when its busiest XCD does.
#pragma once

#include <algorithm>
#include <vector>

namespace onh {

constexpr int kXcds = 8;

// Each XCD gets a contiguous run of the tiles in natural (system-major, row-major) order, runs of
// equal length.
template <typename TileT>
std::vector<TileT> xcd_order(const std::vector<TileT>& nat) {
  const int T = (int)nat.size(), X = kXcds;
  std::vector<TileT> out(T);
  std::vector<int> start(X + 1, 0);
  for (int x = 0; x < X; ++x) start[x + 1] = start[x] + (T > x ? (T - x + X - 1) / X : 0);
  for (int b = 0; b < T; ++b) out[b] = nat[start[b % X] + b / X];
  return out;
}

// The order of a streaming launch. A tile's work grows with its system's levels, so equal runs of
// tiles are unequal work when the systems of a launch differ in depth (the 64-system batch: 5.4 % above
// the mean on the busiest XCD). The tiles, in natural order, are cut instead into 8 contiguous runs of
// equal work (levels + 4 per tile for the block's set-up), each tile going to the run its midpoint
// falls in; the shorter runs are padded with `empty` tiles, which the kernels skip. Inside a run the
// deepest systems' tiles go first (stable: a system's tiles keep their order), so the last blocks to
// start are the shortest. Equal depths give contiguous runs of equal length (within one tile), as
// xcd_order does. Lists of fewer than min_tiles tiles keep xcd_order. depth(t): the levels of tile t's system.
template <typename TileT, typename Depth>
std::vector<TileT> balanced_order(const std::vector<TileT>& nat, Depth depth, int min_tiles, const TileT& empty) {
  const int T = (int)nat.size(), X = kXcds;
  if (T < min_tiles) return xcd_order(nat);
  auto work = [&](const TileT& t) { return (double)depth(t) + 4.0; };
  double total = 0.0;
  for (const TileT& t : nat) total += work(t);
  std::vector<std::vector<TileT>> run(X);
  double acc = 0.0;
  int x = 0;
  for (const TileT& t : nat) {
    const double mid = acc + 0.5 * work(t);
    while (x < X - 1 && mid > total * (x + 1) / X) ++x;
    run[x].push_back(t);
    acc += work(t);
  }
  size_t len = 0;
  for (auto& r : run) {
    std::stable_sort(r.begin(), r.end(), [&](const TileT& a, const TileT& b) { return depth(a) > depth(b); });
    len = std::max(len, r.size());
  }
  std::vector<TileT> out(len * X, empty);
  for (int r = 0; r < X; ++r)
    for (size_t k = 0; k < run[r].size(); ++k) out[k * X + r] = run[r][k];
  return out;
}

// The blocks of the paired column update, from a launch's tiles in natural order (each system's
// tiles row-major, bx fastest). A block covers two tiles: (bx, by) with bx even and its right
// neighbour (bx + 1, by) (local = 0). Where a system's rows hold an odd number of tiles, the last tile
// of row-block by (even) is paired with the last tile of row-block by + 1 below it (local = 1), so two
// partly filled blocks become one; a last tile with no partner runs alone (local = 0, no right
// neighbour). Every tile is covered exactly once.
template <typename TileT>
std::vector<TileT> pair_tiles(const std::vector<TileT>& nat) {
  std::vector<TileT> out;
  const int n = (int)nat.size();
  int a = 0;
  while (a < n) {
    int e = a, tx = 0, ty = 0;  // the system's tiles are nat[a, e)
    while (e < n && nat[e].sys == nat[a].sys) {
      tx = std::max(tx, nat[e].bx + 1);
      ty = std::max(ty, nat[e].by + 1);
      ++e;
    }
    for (int k = a; k < e; ++k) {
      TileT t = nat[k];
      if (t.bx % 2 != 0) continue;
      t.local = 0;
      if (t.bx + 1 == tx) {  // the last tile of an odd row
        if (t.by % 2 != 0) continue;  // paired with the one above
        if (t.by + 1 < ty) t.local = 1;
      }
      out.push_back(t);
    }
    a = e;
  }
  return out;
}

}  // namespace onh
