// nh_sor_tb.h -- the red-black SOR sweeps of TRONEU (troneu.f90:104-131) with temporal blocking: one
// launch performs up to T half-sweeps on an output tile of (RX-2T) x (RY-2T) cells, from a region of
// RX x RY cells staged in LDS. Each half-sweep invalidates one more ring of the region, so after T
// half-sweeps the inner tile still holds exactly the values of the sequential sweeps.
//
// Virtual coordinates. Region cells carry virtual coordinates (vx, vy) that may lie outside 1..IE,
// 1..JE. A cyclic direction maps them periodically onto the interior, 2 + mod(v - 2, E - 2): the halo
// cell 1 becomes cell E-1 and the halo cell E becomes cell 2, which is what bounds_exch would have copied
// there after the previous half-sweep. A non-cyclic direction keeps 1..E (the halo rows and columns
// are never updated and keep their values) and marks everything outside as absent. With this map the
// stored halo of a cyclic direction is never read during the sweeps; one bounds_exch after the last
// launch writes it, as the last exchange of the source does.
//
// Jacobi order within a half-sweep. Every update of a half-sweep reads the values from before the
// half-sweep (all reads, barrier, all writes). Inside the domain an active cell's four neighbours
// have the other colour and are not updated in the same half-sweep, so this is the sequential order's
// result; across a cyclic seam with an odd period a neighbour can share the colour, and the source then
// reads the halo copy, which also holds the value from before the half-sweep.
//
// Work split. A thread owns NP vertical pairs (x, 2p) and (x, 2p+1), one cell of each colour, so every
// lane updates one cell per half-sweep; consecutive lanes own consecutive x, so the region loads and
// the tile stores are contiguous rows. The LDS holds the region split by colour: plane (x + y) & 1,
// row y >> 1, column x, padded by a row and a word on either side, so the four neighbours of a cell lie
// in the other plane, consecutive lanes read consecutive words, and a lane's reads for all its pairs
// are one base address plus constant offsets. B1O and SORPAR*FF of every cell sit beside them in LDS;
// the four face coefficients of a thread's cells stay in registers for the whole launch, ordered by
// the colour that updates them, which keeps a 512-thread block within 256 VGPRs at five pairs per thread.
//
// Regular and seam kernels. Unless a cyclic direction has an odd period, the mapped parity equals the
// virtual one everywhere, so a lane's colour-C cells are all cell s0 ^ C of their pairs,
// s0 = (vx0 + x + vy0) & 1, and the kernel (SEAM = false) runs a branch-free path: the region's edge
// cells compute from the padding and are not written back. With an odd cyclic period (SEAM = true)
// every pair checks which of its cells the colour updates, both or none near the seam.
//
// Everything here is __host__ __device__ so that the same code can be stepped thread by thread on
// the host. FP64, expressions left to right as the Fortran source writes them (no contraction):
// S = B1O + UF(I)Z(I+1) + UF(I-1)Z(I-1) + VF(J)Z(J+1) + VF(J-1)Z(J-1), Z = SORPAR*FF*S + SORPAI*Z.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace onh {

struct TbGeom {
  int IE, JE, IS, JS, NU;  // p-points (IE,JE), u-points (IS:IE,JE) with NU = IE-IS+1, v-points (IE,JS:JE)
  int cyc_x, cyc_y;
  int tiles_x;  // output tiles along i
};

// virtual coordinate -> array coordinate, 0 = absent
__host__ __device__ inline int tb_map(int v, int E, int cyc) {
  if (cyc) {
    const int P = E - 2;
    int r = (v - 2) % P;
    if (r < 0) r += P;
    return 2 + r;
  }
  return (v >= 1 && v <= E) ? v : 0;
}

// Block -> tile so that each of the 8 XCDs (which receive blocks round robin) gets a contiguous run of
// tiles, whose shared region halos then meet in the same L2. A bijection on 0..nb-1.
__host__ __device__ inline int tb_tile(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7, x = b & 7, slot = b >> 3;
  return x * per + (x < rem ? x : rem) + slot;
}

struct TbCo {
  double ue, uw, vn, vs;  // UF(I), UF(I-1), VF(J), VF(J-1)
};

template <int T, int RX, int RY, int NT, bool SEAM>
struct SorTb {
  static constexpr int PR = RY / 2;        // pair rows
  static constexpr int PL = RX * PR;       // cells per colour plane = pairs per region
  static constexpr int NP = PL / NT;       // pairs per thread
  static constexpr int DP = NT / RX;       // pair rows between consecutive pairs of a thread
  static constexpr int PAD = RX + 1;       // padding words before and after each plane
  static constexpr int PS = PL + 2 * PAD;  // plane stride
  static constexpr int BW = 2 * PS;        // start of the (B1O, SORPAR*FF) pairs, same plane order, no padding
  static constexpr int LDS = 2 * PS + 4 * PL;  // doubles of LDS
  static constexpr int TXo = RX - 2 * T;   // output tile
  static constexpr int TYo = RY - 2 * T;
  static_assert(RY % 2 == 0 && NT % RX == 0 && PL % NT == 0 && NP >= 1 && NP <= 16,
                "region must split into whole pairs per thread, with a fixed column per thread");
  static_assert(TXo >= 1 && TYo >= 1, "halo wider than the region");
  static_assert(NT % 64 == 0, "whole wavefronts");

  struct Regs {
    TbCo c[2][NP];     // c[C][k]: coefficients of cell s0 ^ C of pair k (the one colour C updates
                       // away from odd-period seams)
    double z[NP][2];   // the thread's own cells (kept current by the regular path)
    double n0[NP], n1[NP];
    uint32_t m0, m1;   // bit 2k+s: cell s of pair k is updated in colour 0 / colour 1
    int s0;
  };

  __host__ __device__ static int at(int x, int y) { return PAD + ((x + y) & 1) * PS + (y >> 1) * RX + x; }
  // index of the (B1O, SORPAR*FF) pair of cell (x, y)
  __host__ __device__ static int bw(int x, int y) { return BW + 2 * (((x + y) & 1) * PL + (y >> 1) * RX + x); }

  static __host__ __device__ size_t p2(const TbGeom& g, int i, int j) {
    return (size_t)(i - 1) + (size_t)g.IE * (size_t)(j - 1);
  }
  static __host__ __device__ size_t u2(const TbGeom& g, int i, int j) {
    return (size_t)(i - g.IS) + (size_t)g.NU * (size_t)(j - 1);
  }
  static __host__ __device__ size_t v2(const TbGeom& g, int i, int j) {
    return (size_t)(i - 1) + (size_t)g.IE * (size_t)(j - g.JS);
  }

  // Region origin (virtual coordinates of region cell (0, 0)) of tile t.
  __host__ __device__ static void origin(const TbGeom& g, int t, int& vx0, int& vy0) {
    vx0 = 2 + (t % g.tiles_x) * TXo - T;
    vy0 = 2 + (t / g.tiles_x) * TYo - T;
  }

  __host__ __device__ static void load(int tid, int vx0, int vy0, const TbGeom& g, const double* __restrict__ zin,
                                       const double* __restrict__ b1o, const double* __restrict__ ff,
                                       const double* __restrict__ uf, const double* __restrict__ vf, double sorpar,
                                       double* lz, Regs& r) {
    const int x = tid % RX;
    const int mi = tb_map(vx0 + x, g.IE, g.cyc_x);
    r.m0 = 0u;
    r.m1 = 0u;
    r.s0 = (vx0 + x + vy0) & 1;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int p = tid / RX + DP * k;
      TbCo co[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int y = 2 * p + s;
        const int mj = tb_map(vy0 + y, g.JE, g.cyc_y);
        const bool in = mi != 0 && mj != 0;
        const bool upd = in && x >= 1 && x <= RX - 2 && y >= 1 && y <= RY - 2 && mi >= 2 && mi <= g.IE - 1 &&
                         mj >= 2 && mj <= g.JE - 1;
        // every load is issued, from cell (2, 2) when the cell is absent or not updated, so that the
        // thread's loads go out back to back
        const int zi = in ? mi : 2, zj = in ? mj : 2, ci = upd ? mi : 2, cj = upd ? mj : 2;
        const double zv = zin[p2(g, zi, zj)];
        const double b = b1o[p2(g, ci, cj)], f = ff[p2(g, ci, cj)];
        const double ue = uf[u2(g, ci, cj)], uw = uf[u2(g, ci - 1, cj)];
        const double vn = vf[v2(g, ci, cj)], vs = vf[v2(g, ci, cj - 1)];
        r.z[k][s] = in ? zv : 0.0;
        lz[at(x, y)] = r.z[k][s];
        // SORPAR*FF(I,J) is the first product of the source's expression
        lz[bw(x, y)] = upd ? b : 0.0;
        lz[bw(x, y) + 1] = upd ? sorpar * f : 0.0;
        co[s] = upd ? TbCo{ue, uw, vn, vs} : TbCo{0.0, 0.0, 0.0, 0.0};
        const uint32_t bit = upd ? 1u << (2 * k + s) : 0u;
        if (((mi + mj) & 1) == 0) r.m0 |= bit;
        else r.m1 |= bit;
      }
      r.c[0][k] = r.s0 ? co[1] : co[0];
      r.c[1][k] = r.s0 ? co[0] : co[1];
    }
    // the padding, read by the region's edge cells on the regular path (their results are discarded)
    for (int i = tid; i < 2 * PAD; i += NT) {
      const int o = i < PAD ? i : PS - 2 * PAD + i;
      lz[o] = 0.0;
      lz[PS + o] = 0.0;
    }
  }

  __host__ __device__ static double cell(const TbCo& c, double b, double w, double zE, double zW, double zN, double zS,
                                         double z0, double sorpai) {
    const double s = b + c.ue * zE + c.uw * zW + c.vn * zN + c.vs * zS;
    return w * s + sorpai * z0;
  }

  // New values of this thread's cells of colour C, from the region as it stands (before the writes).
  template <int C>
  __host__ __device__ static void compute(int tid, double sorpai, const double* lz, Regs& r) {
    const int x = tid % RX, p0 = tid / RX;
    if (!SEAM) {
      // The active cell of every pair is cell s; its pair partner (cell 1-s, the neighbour above for
      // s = 0 and below for s = 1) and its own value are in registers, the other vertical neighbour
      // (in the pair row above or below) and the two horizontal ones come from the other plane.
      const int s = r.s0 ^ C, own = (x + s) & 1;
      const double* lo = lz + PAD + (1 - own) * PS + p0 * RX + x;  // the other plane at (x, pair row p0)
      const double* lb = lz + BW + 2 * (own * PL + p0 * RX + x);   // this cell's (B1O, SORPAR*FF)
      const int dv = s ? RX : -RX;  // the pair row of the vertical neighbour outside the pair
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int o = k * DP * RX;
        const double zE = lo[o + 1], zW = lo[o - 1], zV = lo[o + dv];
        const double b = lb[2 * o], w = lb[2 * o + 1];
        const double z0 = s ? r.z[k][1] : r.z[k][0], zP = s ? r.z[k][0] : r.z[k][1];
        const double v = cell(r.c[C][k], b, w, zE, zW, s ? zV : zP, s ? zP : zV, z0, sorpai);
        r.n0[k] = v;
        // only this thread reads its registers, and the partner is not updated in this colour
        if ((C ? r.m1 : r.m0) >> (2 * k + s) & 1u) {
          if (s) r.z[k][1] = v;
          else r.z[k][0] = v;
        }
      }
      return;
    }
    const uint32_t m = C ? r.m1 : r.m0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const uint32_t a = (m >> (2 * k)) & 3u;
      const int p = p0 + DP * k;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (a & (1u << s)) {
          const int y = 2 * p + s;
          // coefficient slot of cell s: C' with s0 ^ C' == s
          const TbCo& c = (r.s0 ^ s) ? r.c[1][k] : r.c[0][k];
          const double v = cell(c, lz[bw(x, y)], lz[bw(x, y) + 1], lz[at(x + 1, y)], lz[at(x - 1, y)],
                                lz[at(x, y + 1)], lz[at(x, y - 1)], lz[at(x, y)], sorpai);
          if (s == 0) r.n0[k] = v;
          else r.n1[k] = v;
        }
      }
    }
  }

  template <int C>
  __host__ __device__ static void write(int tid, double* lz, const Regs& r) {
    const int x = tid % RX, p0 = tid / RX;
    const uint32_t m = C ? r.m1 : r.m0;
    if (!SEAM) {
      const int s = r.s0 ^ C, own = (x + s) & 1;
      double* ls = lz + PAD + own * PS + p0 * RX + x;
#pragma unroll
      for (int k = 0; k < NP; ++k)
        if ((m >> (2 * k + s)) & 1u) ls[k * DP * RX] = r.n0[k];
      return;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const uint32_t a = (m >> (2 * k)) & 3u;
      const int p = p0 + DP * k;
      if (a & 1u) lz[at(x, 2 * p)] = r.n0[k];
      if (a & 2u) lz[at(x, 2 * p + 1)] = r.n1[k];
    }
  }

  // The inner tile back to the array (interior cells only; tiles at the far edges are clipped).
  __host__ __device__ static void store(int tid, int vx0, int vy0, const TbGeom& g, const double* lz,
                                        double* __restrict__ zout) {
    const int x = tid % RX, vx = vx0 + x;
    if (x < T || x >= T + TXo || vx > g.IE - 1) return;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int p = tid / RX + DP * k;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int y = 2 * p + s, vy = vy0 + y;
        if (y >= T && y < T + TYo && vy <= g.JE - 1) zout[p2(g, vx, vy)] = lz[at(x, y)];
      }
    }
  }
};

// The half-sweeps per launch that are compiled, with their regions: T, RX, RY, threads (pairs per
// thread = RX*RY/2/threads).
#define ORCTM_SOR_TB_VARIANTS(X) \
  X(4, 64, 32, 256)              \
  X(6, 64, 80, 512)              \
  X(7, 64, 80, 512)              \
  X(8, 64, 64, 512)              \
  X(12, 64, 64, 512)

}  // namespace onh
