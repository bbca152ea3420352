// nh_bot.hip -- the 2-D bottom-pressure solve of the nonhydrostatic step, ORCTM's CG2D4NONHY option
// (ocvtot_nh.f90:64-147, mo_petsc.f90:147-208, 252-374, 401-538; off in the shipped Makefile), on the
// device, for one GPU holding the whole domain.
//
// Per step, after DIVG:
//   k_bot_extract   BOTGU, BOTGV, BOTGW: the tendency at the bottom wet level of every column, as the
//                   source's K loop assigns it (one thread per column walks K)
//   k_bot_rhs       DIVG_bot_surf = bottom divergence - DIVG(:,:,1)*WETO on the interior
//   2-D solve       compute_matrix_2d's operator on the (IE-2) x (JE-2) interior box: the box-edge
//                   rows (diagonal only, Neumann = .FALSE.) directly, then CG on the remaining rows
//                   scaled by S = DLXP*DLYP, preconditioned by one multigrid V-cycle (2 x 2 aggregates,
//                   Galerkin operators, damped Jacobi smoothing, a dense solve on the last level),
//                   warm-started from the previous solution as KSPSetInitialGuessNonzero does
//                   (mo_petsc.f90:165), to ||b - A2 p|| <= rtol ||b|| checked every kBotChunk
//                   iterations; block sums in fixed order and V-cycle kernels whose cell values do
//                   not depend on the blocking, so the iterates are the same bits on every run
//   k2_scatter      PNH_bot on the interior, then bounds_exch(PNH_bot) (ocvtot_nh.f90:146)
//   k_bot_correct   ComputeRHS's corrections of DIVG by PNH_bot at dry neighbours and the bottom face
// FP64 throughout; the bottom values, DIVG_bot_surf and the corrections evaluate each expression left
// to right as the source writes it (no contraction), so they are the oracle's bits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>
#include <vector>

#include "../../include/orctm_nh.h"
#include "nh_bot.h"

namespace onh {

namespace {

constexpr int BB = 256;  // threads per block

struct BDims {
  int IE, JE, KE, IS, JS, NU, NV, nx, ny;
  __device__ size_t p2(int i, int j) const { return (size_t)(i - 1) + (size_t)IE * (size_t)(j - 1); }
  __device__ size_t p3(int i, int j, int k) const {
    return (size_t)(i - 1) + (size_t)IE * ((size_t)(j - 1) + (size_t)JE * (size_t)(k - 1));
  }
  __device__ size_t u2(int i, int j) const { return (size_t)(i - IS) + (size_t)NU * (size_t)(j - 1); }
  __device__ size_t u3(int i, int j, int k) const {
    return (size_t)(i - IS) + (size_t)NU * ((size_t)(j - 1) + (size_t)JE * (size_t)(k - 1));
  }
  __device__ size_t v2(int i, int j) const { return (size_t)(i - 1) + (size_t)IE * (size_t)(j - JS); }
  __device__ size_t v3(int i, int j, int k) const {
    return (size_t)(i - 1) + (size_t)IE * ((size_t)(j - JS) + (size_t)NV * (size_t)(k - 1));
  }
  __device__ bool inner(int c) const {
    const int i = c % nx, j = c / nx;
    return i > 0 && j > 0 && i < nx - 1 && j < ny - 1;
  }
};

BDims dims(const BotState& b) {
  return BDims{b.IE, b.JE, b.KE, b.IS, b.JS, b.IE - b.IS + 1, b.JE - b.JS + 1, b.nx, b.ny};
}

inline unsigned nblocks(long long n) { return (unsigned)((n + BB - 1) / BB); }

enum { S_RZ = 0, S_DQ, S_ALPHA, S_BETA, S_RR, S_BB, S_FLAG, S_N };

// Block sums of up to three values per thread into part[3*block + m], in a fixed tree order. Every
// thread of the block must call it.
__device__ void block_sums(double v0, double v1, double v2, double* part) {
  __shared__ double sh[3][BB];
  const int t = threadIdx.x;
  sh[0][t] = v0;
  sh[1][t] = v1;
  sh[2][t] = v2;
  __syncthreads();
  for (int s = BB / 2; s > 0; s >>= 1) {
    if (t < s) {
      sh[0][t] += sh[0][t + s];
      sh[1][t] += sh[1][t + s];
      sh[2][t] += sh[2][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[3 * blockIdx.x] = sh[0][0];
    part[3 * blockIdx.x + 1] = sh[1][0];
    part[3 * blockIdx.x + 2] = sh[2][0];
  }
}

// mo_petsc.f90:415-527 (Neumann = .FALSE.): the 2-D operator on the interior box, and S = DLXP*DLYP.
__global__ __launch_bounds__(BB) void k_bot_coef(BDims g, const double* __restrict__ dlxu,
                                                 const double* __restrict__ dlyv, const double* __restrict__ dlxp,
                                                 const double* __restrict__ dlyp, double* __restrict__ coef,
                                                 double* __restrict__ S) {
  const int c = blockIdx.x * BB + threadIdx.x;
  const int n = g.nx * g.ny;
  if (c >= n) return;
  const int i = 2 + c % g.nx, j = 2 + c / g.nx;
  const double xuw = 1.0 / dlxu[g.u2(i - 1, j)], xue = 1.0 / dlxu[g.u2(i, j)], xp = 1.0 / dlxp[g.p2(i, j)];
  const double yvn = 1.0 / dlyv[g.v2(i, j - 1)], yvs = 1.0 / dlyv[g.v2(i, j)], yp = 1.0 / dlyp[g.p2(i, j)];
  const double X_W = xuw * xp, X_C = (xuw + xue) * xp, X_E = xue * xp;
  const double Y_N = yvn * yp, Y_C = (yvn + yvs) * yp, Y_S = yvs * yp;
  const bool edge = i == 2 || j == 2 || i == g.IE - 1 || j == g.JE - 1;
  coef[c] = -(X_C + Y_C);
  coef[(size_t)n + c] = edge ? 0.0 : X_W;
  coef[2 * (size_t)n + c] = edge ? 0.0 : X_E;
  coef[3 * (size_t)n + c] = edge ? 0.0 : Y_N;
  coef[4 * (size_t)n + c] = edge ? 0.0 : Y_S;
  S[c] = dlxp[g.p2(i, j)] * dlyp[g.p2(i, j)];
}

// ocvtot_nh.f90:64-124: one thread per u-, v- or p-column walks K and applies the source's assignments
// in order (K = 1 with the top face dry: 0; K = KE with the bottom face wet: G(KE); otherwise a wet-to-dry
// step between K-1 and K: G(K-1)); with none the stored value stays.
__global__ __launch_bounds__(BB) void k_bot_extract(BDims g, OcvArrays a, double* __restrict__ botgu,
                                                    double* __restrict__ botgv, double* __restrict__ botgw) {
  const int n = blockIdx.x * BB + threadIdx.x;
  const int nu = g.NU * g.JE, nv = g.IE * g.NV, nw = g.IE * g.JE;
  const double dt = a.dt;
  if (n < nu) {
    const int I = g.IS + n % g.NU, J = 1 + n / g.NU;
    double v = botgu[g.u2(I, J)];
    for (int K = 1; K <= g.KE; ++K) {
      const double am = a.amsuo[g.u3(I, J, K)];
      if (K == 1) {
        if (am == 0.0) v = 0.0;
      } else if (K == g.KE) {
        if (am != 0.0) v = (a.uoo[g.u3(I, J, K)] - a.uoo_1[g.u3(I, J, K)]) / dt;
      } else if (a.amsuo[g.u3(I, J, K - 1)] != 0.0 && am == 0.0) {
        v = (a.uoo[g.u3(I, J, K - 1)] - a.uoo_1[g.u3(I, J, K - 1)]) / dt;
      }
    }
    botgu[g.u2(I, J)] = v;
  } else if (n < nu + nv) {
    const int m = n - nu;
    const int I = 1 + m % g.IE, J = g.JS + m / g.IE;
    double v = botgv[g.v2(I, J)];
    for (int K = 1; K <= g.KE; ++K) {
      const double am = a.amsue[g.v3(I, J, K)];
      if (K == 1) {
        if (am == 0.0) v = 0.0;
      } else if (K == g.KE) {
        if (am != 0.0) v = (a.voe[g.v3(I, J, K)] - a.voe_1[g.v3(I, J, K)]) / dt;
      } else if (a.amsue[g.v3(I, J, K - 1)] != 0.0 && am == 0.0) {
        v = (a.voe[g.v3(I, J, K - 1)] - a.voe_1[g.v3(I, J, K - 1)]) / dt;
      }
    }
    botgv[g.v2(I, J)] = v;
  } else if (n < nu + nv + nw) {
    const int m = n - nu - nv;
    const int I = 1 + m % g.IE, J = 1 + m / g.IE;
    double v = botgw[g.p2(I, J)];
    for (int K = 1; K <= g.KE; ++K) {
      const double w = a.weto[g.p3(I, J, K)];
      if (K == 1) {
        if (w == 0.0) v = 0.0;
      } else if (K == g.KE) {
        if (w != 0.0) v = (a.wno[g.p3(I, J, K)] - a.wno_1[g.p3(I, J, K)]) / dt / a.ddpo[g.p3(I, J, K)];
      } else if (a.weto[g.p3(I, J, K - 1)] != 0.0 && w == 0.0) {
        v = (a.wno[g.p3(I, J, K - 1)] - a.wno_1[g.p3(I, J, K - 1)]) / dt / a.ddpo[g.p3(I, J, K - 1)];
      }
    }
    botgw[g.p2(I, J)] = v;
  }
}

// ocvtot_nh.f90:127-144 on the interior: DIVBOTGU = (BOTGU(I)-BOTGU(I-1))/DLXP, DIVBOTGV =
// (BOTGV(J-1)-BOTGV(J))/DLYP, DIVBOTGW = BOTGW, DIVG_bot_surf = (sum) - DIVG(I,J,1)*WETO(I,J,1); also the
// 2-D right-hand side (ComputeRHS_2d, mo_petsc.f90:178-182).
__global__ __launch_bounds__(BB) void k_bot_rhs(BDims g, OcvArrays a, const double* __restrict__ botgu,
                                                const double* __restrict__ botgv, const double* __restrict__ botgw,
                                                const double* __restrict__ divg, double* __restrict__ dbs,
                                                double* __restrict__ b) {
  const int c = blockIdx.x * BB + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int I = 2 + c % g.nx, J = 2 + c / g.nx;
  const double du = (botgu[g.u2(I, J)] - botgu[g.u2(I - 1, J)]) / a.dlxp[g.p2(I, J)];
  const double dv = (botgv[g.v2(I, J - 1)] - botgv[g.v2(I, J)]) / a.dlyp[g.p2(I, J)];
  const double dw = botgw[g.p2(I, J)];
  const double v = (du + dv + dw) - divg[g.p3(I, J, 1)] * a.weto[g.p3(I, J, 1)];
  dbs[g.p2(I, J)] = v;
  b[c] = v;
}

// (A2 v)_c on an inner row, summed north, west, centre, east, south
__device__ inline double apply_inner(const BDims& g, const double* coef, int c, double vn, double vw, double vc,
                                     double ve, double vs) {
  const size_t n = (size_t)g.nx * g.ny;
  double s = 0.0;
  s += coef[3 * n + c] * vn;
  s += coef[n + c] * vw;
  s += coef[c] * vc;
  s += coef[2 * n + c] * ve;
  s += coef[4 * n + c] * vs;
  return s;
}

// The box-edge rows directly (x = b / A_cc); inner x kept (warm start) or zeroed; search directions zeroed.
__global__ __launch_bounds__(BB) void k2_start(BDims g, const double* __restrict__ coef,
                                               const double* __restrict__ b, double* __restrict__ x,
                                               double* __restrict__ d0, double* __restrict__ d1, int cold) {
  const int c = blockIdx.x * BB + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  if (!g.inner(c)) x[c] = b[c] / coef[c];
  else if (cold) x[c] = 0.0;
  d0[c] = 0.0;
  d1[c] = 0.0;
}

// t = b - A2 x on every row; partial sums of t^2 and b^2; with init also r = -S t on the inner rows
// (0 on the edge), which the V-cycle then preconditions.
__global__ __launch_bounds__(BB) void k2_resid(BDims g, const double* __restrict__ coef,
                                               const double* __restrict__ S, const double* __restrict__ b,
                                               const double* __restrict__ x, double* __restrict__ r,
                                               double* __restrict__ part, int init) {
  const int c = blockIdx.x * BB + threadIdx.x;
  const int n = g.nx * g.ny;
  double tt = 0.0, bb = 0.0;
  if (c < n) {
    double ax;
    const bool in = g.inner(c);
    if (in) ax = apply_inner(g, coef, c, x[c - g.nx], x[c - 1], x[c], x[c + 1], x[c + g.nx]);
    else ax = coef[c] * x[c];
    const double t = b[c] - ax;
    tt = t * t;
    bb = b[c] * b[c];
    if (init) r[c] = in ? -S[c] * t : 0.0;
  }
  block_sums(tt, bb, 0.0, part);
}

// d = z + beta d_old on the inner rows (zero on the edge), q = -S A2 d, partial sum of d q. The
// neighbours' new d are formed here from their z and d_old, so d_old stays read-only.
__global__ __launch_bounds__(BB) void k2_dirmv(BDims g, const double* __restrict__ coef,
                                               const double* __restrict__ S, const double* __restrict__ z,
                                               const double* __restrict__ dold, double* __restrict__ dnew,
                                               double* __restrict__ q, double* __restrict__ part,
                                               const double* __restrict__ scal) {
  const int c = blockIdx.x * BB + threadIdx.x;
  const int n = g.nx * g.ny;
  const double beta = scal[S_BETA];
  double dq = 0.0;
  if (c < n) {
    if (g.inner(c)) {
      auto dn = [&](int cc) { return g.inner(cc) ? z[cc] + beta * dold[cc] : 0.0; };
      const double dc = z[c] + beta * dold[c];
      const double ad = apply_inner(g, coef, c, dn(c - g.nx), dn(c - 1), dc, dn(c + 1), dn(c + g.nx));
      const double qc = -S[c] * ad;
      dnew[c] = dc;
      q[c] = qc;
      dq = dc * qc;
    } else {
      dnew[c] = 0.0;
      q[c] = 0.0;
    }
  }
  block_sums(dq, 0.0, 0.0, part);
}

// x += alpha d, r -= alpha q on the inner rows (the V-cycle then forms z and r.z).
__global__ __launch_bounds__(BB) void k2_update(BDims g, double* __restrict__ x, const double* __restrict__ d,
                                                double* __restrict__ r, const double* __restrict__ q,
                                                const double* __restrict__ scal) {
  const int c = blockIdx.x * BB + threadIdx.x;
  const double alpha = scal[S_ALPHA];
  if (c < g.nx * g.ny && g.inner(c)) {
    x[c] += alpha * d[c];
    r[c] = r[c] - alpha * q[c];
  }
}

// ---- the multigrid V-cycle (the host issues the levels' kernels; the coarsest levels run in one block) --
// The operations on a level's cells (K the level's operator, unknowns the cells with MgDev::active):
//   first sweep from x = 0:  x = omega r / K_cc on the unknowns, 0 elsewhere
//   sweep:                   y = x + omega (r - K x) / K_cc on the unknowns, 0 elsewhere
//   restriction:             rc = the sum of r - K x over an aggregate's unknowns, children in the order
//                            (0,0), (1,0), (0,1), (1,1)
//   coarse correction:       x += kAlpha xc of the cell's aggregate, on the unknowns
//   dense solve:             x = inv r over the last level's unknowns (numbered row by row), 0 elsewhere
// K x is MgDev::kx (centre, west, east, north, south). Whichever kernel computes a value computes it by
// these operations in this order, so the V-cycle's results do not depend on how it is launched.

constexpr double kOmega = 2.0 / 3.0;  // damped Jacobi
constexpr double kAlpha = 1.8;        // scale of every coarse-grid correction
constexpr int kNu = 2;                // Jacobi sweeps before and after each coarse correction

struct MgDev {
  int a, b, o;
  const double* k;
  __device__ bool active(int i, int j) const { return i >= o && j >= o && i < a - o && j < b - o; }
  __device__ double kx(const double* __restrict__ x, int i, int j, int c) const {
    const size_t n = (size_t)a * b;
    double s = k[c] * x[c];
    if (i > 0) s += k[n + c] * x[c - 1];
    if (i < a - 1) s += k[2 * n + c] * x[c + 1];
    if (j > 0) s += k[3 * n + c] * x[c - a];
    if (j < b - 1) s += k[4 * n + c] * x[c + a];
    return s;
  }
};

// x = inv r over the level's unknowns (numbered row by row), 0 elsewhere; one block
__global__ __launch_bounds__(BB) void k_mg_dense(MgDev L, int n, const double* __restrict__ inv,
                                                 const double* __restrict__ r, double* __restrict__ x) {
  const int wa = L.a - 2 * L.o;
  for (int c = threadIdx.x; c < L.a * L.b; c += BB) {
    const int i = c % L.a, j = c / L.a;
    double s = 0.0;
    if (L.active(i, j)) {
      const int t = (j - L.o) * wa + (i - L.o);
      for (int q = 0; q < n; ++q) {
        const int cq = (q / wa + L.o) * L.a + (q % wa + L.o);
        s += inv[(size_t)t * n + q] * r[cq];
      }
    }
    x[c] = s;
  }
}

// the block sums of r z over the box (when the last level is the box itself)
__global__ __launch_bounds__(BB) void k_mg_dot(int n, const double* __restrict__ r, const double* __restrict__ z,
                                               double* __restrict__ part) {
  const int c = blockIdx.x * BB + threadIdx.x;
  block_sums(c < n ? r[c] * z[c] : 0.0, 0.0, 0.0, part);
}

// ---- a level's way down in one launch -------------------------------------------------------------
// With kNu = 2 the way down through a level is the first sweep from x = 0, one more sweep (y) and the
// restriction of y. k_mg_down does all three for a block of CTX x CTY aggregates: r and K of its cells and a ring of 2
// go to LDS, the first sweep (pointwise) runs on all of them, the second on the cells and a ring of 1,
// then the residuals of the cells and their aggregate sums follow. The first and last blocks of a row
// or column of blocks also take the level's cells outside every aggregate. Only y goes to memory: x
// is overwritten, unread, by the first sweep on the way up. Each value is computed by the same
// operations in the same order whichever block computes it, so the results do not depend on the
// blocking.
constexpr int CTX = 8, CTY = 4;  // aggregates per block (the fastest of 4-16 x 2-8 on a 1024 x 768 box)
constexpr int EXW = 2 * CTX + 6, EXH = 2 * CTY + 6;  // LDS extents: the block's cells (at most 2 CT + 2) and the ring

__device__ __forceinline__ int down_tiles_x(const MgDev& C) { return (C.a + CTX - 1) / CTX; }

__global__ __launch_bounds__(BB) void k_mg_down(MgDev F, MgDev C, const double* __restrict__ r,
                                                double* __restrict__ x1g, double* __restrict__ rc) {
  static_assert(kNu == 2, "the fused way down runs exactly two sweeps");
  __shared__ double sk[5][EXH][EXW], sr[EXH][EXW], s0[EXH][EXW], s1[EXH][EXW], se[EXH][EXW];
  const int ntx = down_tiles_x(C), nty = (C.b + CTY - 1) / CTY;
  const int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  const int ilo = tx == 0 ? 0 : F.o + 2 * CTX * tx, ihi = tx == ntx - 1 ? F.a : F.o + 2 * CTX * (tx + 1);
  const int jlo = ty == 0 ? 0 : F.o + 2 * CTY * ty, jhi = ty == nty - 1 ? F.b : F.o + 2 * CTY * (ty + 1);
  const int ei = ilo - 2, ej = jlo - 2, ew = ihi - ilo + 4, eh = jhi - jlo + 4;
  const size_t n = (size_t)F.a * F.b;
  // K x at LDS cell (li, lj) = level cell (i, j), in MgDev::kx's order
  auto kx = [&](double (*x)[EXW], int li, int lj, int i, int j) {
    double s = sk[0][lj][li] * x[lj][li];
    if (i > 0) s += sk[1][lj][li] * x[lj][li - 1];
    if (i < F.a - 1) s += sk[2][lj][li] * x[lj][li + 1];
    if (j > 0) s += sk[3][lj][li] * x[lj - 1][li];
    if (j < F.b - 1) s += sk[4][lj][li] * x[lj + 1][li];
    return s;
  };
  for (int e = threadIdx.x; e < ew * eh; e += BB) {
    const int li = e % ew, lj = e / ew, i = ei + li, j = ej + lj;
    if (i < 0 || j < 0 || i >= F.a || j >= F.b) continue;
    const size_t c = (size_t)j * F.a + i;
    const double rv = r[c];
    sr[lj][li] = rv;
#pragma unroll
    for (int q = 0; q < 5; ++q) sk[q][lj][li] = F.k[q * n + c];
    s0[lj][li] = F.active(i, j) ? kOmega * rv / sk[0][lj][li] : 0.0;
  }
  __syncthreads();
  const int w1 = ew - 2, h1 = eh - 2;  // the cells and a ring of 1
  for (int e = threadIdx.x; e < w1 * h1; e += BB) {
    const int li = 1 + e % w1, lj = 1 + e / w1, i = ei + li, j = ej + lj;
    if (i < 0 || j < 0 || i >= F.a || j >= F.b) continue;
    double v = 0.0;
    if (F.active(i, j)) v = s0[lj][li] + kOmega * (sr[lj][li] - kx(s0, li, lj, i, j)) / sk[0][lj][li];
    s1[lj][li] = v;
  }
  __syncthreads();
  const int w0 = ew - 4, h0 = eh - 4;  // the block's own cells
  for (int e = threadIdx.x; e < w0 * h0; e += BB) {
    const int li = 2 + e % w0, lj = 2 + e / w0, i = ei + li, j = ej + lj;
    x1g[(size_t)j * F.a + i] = s1[lj][li];
    if (F.active(i, j)) se[lj][li] = sr[lj][li] - kx(s1, li, lj, i, j);
  }
  __syncthreads();
  const int I1 = min(C.a, CTX * (tx + 1)), J1 = min(C.b, CTY * (ty + 1));
  for (int e = threadIdx.x; e < CTX * CTY; e += BB) {
    const int I = CTX * tx + e % CTX, J = CTY * ty + e / CTX;
    if (I >= I1 || J >= J1) continue;
    double s = 0.0;
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        const int i = F.o + 2 * I + di, j = F.o + 2 * J + dj;
        if (F.active(i, j)) s += se[j - ej][i - ei];
      }
    rc[(size_t)J * C.a + I] = s;
  }
}

// A level's way up in one launch: the coarse correction and kNu = 2 sweeps for a block of 2 CTX x 2 CTY
// cells, with r, K and the corrected iterate of its cells and a ring of 2 in LDS, the first sweep on
// the cells and a ring of 1, the second on the cells, written to `out` (not to the iterate read, which
// other blocks still read). On the box, `out` is z and k_mg_dotm forms the r.z block sums after it.
__global__ __launch_bounds__(BB) void k_mg_up(MgDev F, MgDev C, const double* __restrict__ r,
                                              const double* __restrict__ x, const double* __restrict__ xc,
                                              double* __restrict__ out) {
  static_assert(kNu == 2, "the fused way up runs exactly two sweeps");
  constexpr int TW = 2 * CTX, TH = 2 * CTY;
  __shared__ double sk[5][EXH][EXW], sr[EXH][EXW], sa[EXH][EXW], sb[EXH][EXW];
  const int ntx = (F.a + TW - 1) / TW;
  const int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  const int ilo = TW * tx, ihi = min(F.a, ilo + TW), jlo = TH * ty, jhi = min(F.b, jlo + TH);
  const int ei = ilo - 2, ej = jlo - 2, ew = ihi - ilo + 4, eh = jhi - jlo + 4;
  const size_t n = (size_t)F.a * F.b;
  auto kx = [&](double (*v)[EXW], int li, int lj, int i, int j) {
    double s = sk[0][lj][li] * v[lj][li];
    if (i > 0) s += sk[1][lj][li] * v[lj][li - 1];
    if (i < F.a - 1) s += sk[2][lj][li] * v[lj][li + 1];
    if (j > 0) s += sk[3][lj][li] * v[lj - 1][li];
    if (j < F.b - 1) s += sk[4][lj][li] * v[lj + 1][li];
    return s;
  };
  for (int e = threadIdx.x; e < ew * eh; e += BB) {
    const int li = e % ew, lj = e / ew, i = ei + li, j = ej + lj;
    if (i < 0 || j < 0 || i >= F.a || j >= F.b) continue;
    const size_t c = (size_t)j * F.a + i;
    sr[lj][li] = r[c];
#pragma unroll
    for (int q = 0; q < 5; ++q) sk[q][lj][li] = F.k[q * n + c];
    double v = x[c];
    if (F.active(i, j)) v += kAlpha * xc[((j - F.o) / 2) * C.a + (i - F.o) / 2];
    sa[lj][li] = v;
  }
  __syncthreads();
  const int w1 = ew - 2, h1 = eh - 2;
  for (int e = threadIdx.x; e < w1 * h1; e += BB) {
    const int li = 1 + e % w1, lj = 1 + e / w1, i = ei + li, j = ej + lj;
    if (i < 0 || j < 0 || i >= F.a || j >= F.b) continue;
    double v = 0.0;
    if (F.active(i, j)) v = sa[lj][li] + kOmega * (sr[lj][li] - kx(sa, li, lj, i, j)) / sk[0][lj][li];
    sb[lj][li] = v;
  }
  __syncthreads();
  const int w0 = ew - 4, h0 = eh - 4;
  for (int e = threadIdx.x; e < w0 * h0; e += BB) {
    const int li = 2 + e % w0, lj = 2 + e / w0, i = ei + li, j = ej + lj;
    double v = 0.0;
    if (F.active(i, j)) v = sb[lj][li] + kOmega * (sr[lj][li] - kx(sb, li, lj, i, j)) / sk[0][lj][li];
    out[(size_t)j * F.a + i] = v;
  }
}

// the block sums of r z over the box's unknowns (0 elsewhere), in blocks of BB consecutive cells, for
// the k2_reduce that follows (the same block layout as k2_dirmv's sums)
__global__ __launch_bounds__(BB) void k_mg_dotm(MgDev L, const double* __restrict__ r, const double* __restrict__ z,
                                                double* __restrict__ part) {
  const int c = blockIdx.x * BB + threadIdx.x;
  double ry = 0.0;
  if (c < L.a * L.b && L.active(c % L.a, c / L.a)) ry = r[c] * z[c];
  block_sums(ry, 0.0, 0.0, part);
}

// ---- the coarse tail of the V-cycle in one block ---------------------------------------------------
// Below a few hundred cells a level's launches cost far more than its work. The last levels of the
// hierarchy (each at most kTailCells cells, halo included: one cell per thread) therefore run in one
// block: each thread keeps its cell's 5 operator entries in registers, the iterates, right-hand sides
// and residuals of every tail level sit in LDS, and the phases the separate kernels would run are
// separated by block barriers. Every value is computed by the V-cycle's operations in their order (the
// restriction as the cell residuals, then the same 4-child sums), so the results are those of the
// level kernels.
constexpr int kTailCells = 1024;
constexpr int kTailLevels = 4;
constexpr size_t kTailLds = 64 * 1024;  // LDS bytes the tail may use (4 doubles per cell)

struct MgTail {
  int n;  // levels in the tail; the last is the dense level
  int a[kTailLevels], b[kTailLevels], o[kTailLevels];
  const double* k[kTailLevels];
  int dense_n;
  const double* inv;
  const double* r0;  // right-hand side of the first tail level (the restriction from the level above)
  double* x0;        // the first tail level's result, where the prolongation from it reads
};

__global__ __launch_bounds__(kTailCells) void k_mg_tail(MgTail t) {
  extern __shared__ double sm[];
  const int c = threadIdx.x;
  // LDS planes per level: X, Y (the iterate pair), R (right-hand side), E (cell residuals)
  double *X[kTailLevels], *Y[kTailLevels], *R[kTailLevels], *E[kTailLevels];
  double kr[kTailLevels][5];
  {
    double* p = sm;
#pragma unroll
    for (int l = 0; l < kTailLevels; ++l) {
      const int n = l < t.n ? t.a[l] * t.b[l] : 0;
      X[l] = p;
      Y[l] = p + n;
      R[l] = p + 2 * n;
      E[l] = p + 3 * n;
      p += 4 * n;
#pragma unroll
      for (int q = 0; q < 5; ++q) kr[l][q] = (c < n) ? t.k[l][(size_t)q * n + c] : 0.0;
    }
  }
  if (c < t.a[0] * t.b[0]) R[0][c] = t.r0[c];
  __syncthreads();
  auto active = [&](int l, int i, int j) {
    return i >= t.o[l] && j >= t.o[l] && i < t.a[l] - t.o[l] && j < t.b[l] - t.o[l];
  };
  // K x at this thread's cell of level l (MgDev::kx with the entries from registers)
  auto kx = [&](int l, const double* x, int i, int j) {
    const int a = t.a[l];
    double s = kr[l][0] * x[c];
    if (i > 0) s += kr[l][1] * x[c - 1];
    if (i < a - 1) s += kr[l][2] * x[c + 1];
    if (j > 0) s += kr[l][3] * x[c - a];
    if (j < t.b[l] - 1) s += kr[l][4] * x[c + a];
    return s;
  };
  auto jac = [&](int l, const double* x, double* y) {
    if (c < t.a[l] * t.b[l]) {
      const int i = c % t.a[l], j = c / t.a[l];
      double v = 0.0;
      if (active(l, i, j)) v = x[c] + kOmega * (R[l][c] - kx(l, x, i, j)) / kr[l][0];
      y[c] = v;
    }
    __syncthreads();
  };
  double* xf[kTailLevels];
  double* xo[kTailLevels];
  const int L = t.n - 1;
#pragma unroll
  for (int l = 0; l < kTailLevels - 1; ++l) {
    if (l >= L) break;
    const int a = t.a[l], n = a * t.b[l];
    double *xa = X[l], *xb = Y[l];
    if (c < n) {
      const int i = c % a, j = c / a;
      xa[c] = active(l, i, j) ? kOmega * R[l][c] / kr[l][0] : 0.0;
    }
    __syncthreads();
    for (int s = 1; s < kNu; ++s) {
      jac(l, xa, xb);
      double* tmp = xa;
      xa = xb;
      xb = tmp;
    }
    if (c < n) {
      const int i = c % a, j = c / a;
      if (active(l, i, j)) E[l][c] = R[l][c] - kx(l, xa, i, j);
    }
    __syncthreads();
    const int ca = t.a[l + 1];
    if (c < ca * t.b[l + 1]) {
      const int I = c % ca, J = c / ca;
      double s = 0.0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int i = t.o[l] + 2 * I + di, j = t.o[l] + 2 * J + dj;
          if (active(l, i, j)) s += E[l][j * a + i];
        }
      R[l + 1][c] = s;
    }
    __syncthreads();
    xf[l] = xa;
    xo[l] = xb;
  }
  // the dense level: x = inv r over its unknowns (numbered row by row), 0 elsewhere
#pragma unroll
  for (int l = 0; l < kTailLevels; ++l) {
    if (l != L) continue;
    double* x = X[l];
    const int a = t.a[l], o = t.o[l], wa = a - 2 * o;
    if (c < a * t.b[l]) {
      const int i = c % a, j = c / a;
      double s = 0.0;
      if (active(l, i, j)) {
        const int r = (j - o) * wa + (i - o);
        for (int q = 0; q < t.dense_n; ++q) {
          const int cq = (q / wa + o) * a + (q % wa + o);
          s += t.inv[(size_t)r * t.dense_n + q] * R[l][cq];
        }
      }
      x[c] = s;
    }
    __syncthreads();
    xf[l] = x;
  }
#pragma unroll
  for (int l = kTailLevels - 2; l >= 0; --l) {
    if (l >= L) continue;
    const int a = t.a[l], n = a * t.b[l], ca = t.a[l + 1];
    if (c < n) {
      const int i = c % a, j = c / a;
      if (active(l, i, j)) xf[l][c] += kAlpha * xf[l + 1][((j - t.o[l]) / 2) * ca + (i - t.o[l]) / 2];
    }
    __syncthreads();
    double *xa = xf[l], *xb = xo[l];
    for (int s = 1; s <= kNu; ++s) {
      double* out = xb;
      jac(l, xa, out);
      xb = xa;
      xa = out;
    }
    xf[l] = xa;
  }
  if (c < t.a[0] * t.b[0]) t.x0[c] = xf[0][c];
}

enum { R_INIT = 0, R_ALPHA, R_BETA, R_CHECK, R_RZ };

// One block: the block partials summed in a fixed order, then the scalars of the CG recurrence.
__global__ __launch_bounds__(BB) void k2_reduce(const double* __restrict__ part, int nblk, double* __restrict__ scal,
                                                int mode) {
  __shared__ double sh[3][BB];
  const int t = threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = t; k < nblk; k += BB) {
    s0 += part[3 * k];
    s1 += part[3 * k + 1];
    s2 += part[3 * k + 2];
  }
  sh[0][t] = s0;
  sh[1][t] = s1;
  sh[2][t] = s2;
  __syncthreads();
  for (int s = BB / 2; s > 0; s >>= 1) {
    if (t < s) {
      sh[0][t] += sh[0][t + s];
      sh[1][t] += sh[1][t + s];
      sh[2][t] += sh[2][t + s];
    }
    __syncthreads();
  }
  if (t != 0) return;
  if (mode == R_INIT) {
    scal[S_RR] = sh[0][0];
    scal[S_BB] = sh[1][0];
    scal[S_RZ] = sh[2][0];
    scal[S_BETA] = 0.0;
    scal[S_FLAG] = 0.0;
  } else if (mode == R_ALPHA) {
    const double dq = sh[0][0];
    scal[S_DQ] = dq;
    if (dq > 0.0) {
      scal[S_ALPHA] = scal[S_RZ] / dq;
    } else {  // breakdown: stop moving; the host sees the flag at the next check
      scal[S_ALPHA] = 0.0;
      scal[S_FLAG] = 1.0;
    }
  } else if (mode == R_BETA) {
    const double rzn = sh[0][0], rz = scal[S_RZ];
    scal[S_BETA] = rz != 0.0 ? rzn / rz : 0.0;
    scal[S_RZ] = rzn;
  } else if (mode == R_RZ) {
    scal[S_RZ] = sh[0][0];
  } else {
    scal[S_RR] = sh[0][0];
  }
}

__global__ __launch_bounds__(BB) void k2_scatter(BDims g, const double* __restrict__ x, double* __restrict__ pnh_bot) {
  const int c = blockIdx.x * BB + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  pnh_bot[g.p2(2 + c % g.nx, 2 + c / g.nx)] = x[c];
}

// mo_petsc.f90:252-374 with CG2D4NONHY, one thread per interior column walking K. ZW follows
// ComputeRHS's own pre-pass (:252-276: DDPO(K-1)/2 under a dry cell, DZW(KE)-DDPO(KE)/2 at KEP),
// ZP(K) = 1/((1/ZW(K) + 1/ZW(K+1))/2).
__global__ __launch_bounds__(BB) void k_bot_correct(BDims g, const double* __restrict__ dlxu,
                                                    const double* __restrict__ dlyv, const double* __restrict__ dlxp,
                                                    const double* __restrict__ dlyp, const double* __restrict__ dzw,
                                                    const double* __restrict__ ddpo, const double* __restrict__ weto,
                                                    const double* __restrict__ pnh_bot, double* __restrict__ divg) {
  const int c = blockIdx.x * BB + threadIdx.x;
  if (c >= g.nx * g.ny) return;
  const int I = 2 + c % g.nx, J = 2 + c / g.nx;
  const int KE = g.KE, KEP = KE + 1;
  const double xuw = 1.0 / dlxu[g.u2(I - 1, J)], xue = 1.0 / dlxu[g.u2(I, J)], xp = 1.0 / dlxp[g.p2(I, J)];
  const double yvn = 1.0 / dlyv[g.v2(I, J - 1)], yvs = 1.0 / dlyv[g.v2(I, J)], yp = 1.0 / dlyp[g.p2(I, J)];
  const double X_W = xuw * xp, X_C = (xuw + xue) * xp, X_E = xue * xp;
  const double Y_N = yvn * yp, Y_C = (yvn + yvs) * yp, Y_S = yvs * yp;
  const double pc = pnh_bot[g.p2(I, J)];
  auto zw_at = [&](int K) {
    double dd;
    if (K == 1) {
      dd = (ddpo[g.p3(I, J, 1)] == 0.0) ? dzw[0] / 2.0 : ddpo[g.p3(I, J, 1)] / 2.0;
    } else if (K == KEP) {
      dd = (ddpo[g.p3(I, J, KE)] == 0.0) ? dzw[KE - 1] / 2.0 : dzw[KE - 1] - ddpo[g.p3(I, J, KE)] / 2.0;
    } else {
      const double dm = ddpo[g.p3(I, J, K - 1)], d0 = ddpo[g.p3(I, J, K)];
      if (dm == 0.0) dd = (dzw[K - 1] + dzw[K - 2]) / 2.0;
      else if (d0 == 0.0) dd = dm / 2.0;
      else dd = (dm + d0) / 2.0;
    }
    return 1.0 / dd;
  };
  double zwk = zw_at(1);
  for (int K = 1; K <= KE; ++K) {
    const double zwk1 = zw_at(K + 1);
    const double zp = 1.0 / ((1.0 / zwk + 1.0 / zwk1) / 2.0);
    const double Z_C = (zwk + zwk1) * zp, Z_D = zwk1 * zp;
    double Z_U = zwk * zp;
    double dv = divg[g.p3(I, J, K)];
    if (weto[g.p3(I, J, K)] == 0.0) {
      dv = dv + (X_C + Y_C + Z_C) * pc;
    } else {
      if (K == 1) Z_U = 0.0;
      if (K == 1) dv = dv - Z_U * pc;
      else if (weto[g.p3(I, J, K - 1)] == 0.0) dv = dv - Z_U * pc;
      if (weto[g.p3(I, J - 1, K)] == 0.0) dv = dv - Y_N * pnh_bot[g.p2(I, J - 1)];
      if (weto[g.p3(I - 1, J, K)] == 0.0) dv = dv - X_W * pnh_bot[g.p2(I - 1, J)];
      if (weto[g.p3(I + 1, J, K)] == 0.0) dv = dv - X_E * pnh_bot[g.p2(I + 1, J)];
      if (weto[g.p3(I, J + 1, K)] == 0.0) dv = dv - Y_S * pnh_bot[g.p2(I, J + 1)];
      if (K == KE) dv = dv - Z_D * pc;
      else if (weto[g.p3(I, J, K + 1)] == 0.0) dv = dv - Z_D * pc;
    }
    divg[g.p3(I, J, K)] = dv;
    zwk = zwk1;
  }
}

#define BCHK(x)                            \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return e_;       \
  } while (0)

MgDev mgdev(const BotState::MgLevel& L) { return MgDev{L.a, L.b, L.o, L.k}; }

// z = V-cycle(r) on the box, and the block sums of r z into b.part
// The first level of the one-block tail: the coarsest levels (not the box) that each have at most
// kTailCells cells, at most kTailLevels of them and kTailLds bytes of LDS together. The level count
// when there is no tail (the box is the last level).
static int tail_first(const BotState& b) {
  const int L = (int)b.mg.size() - 1;
  int T = L + 1;
  size_t bytes = 0;
  while (T - 1 >= 1) {
    const BotState::MgLevel& lv = b.mg[T - 1];
    const size_t n = (size_t)lv.a * lv.b, need = 4 * sizeof(double) * n;
    if (n > (size_t)kTailCells || L - T + 2 > kTailLevels || bytes + need > kTailLds) break;
    bytes += need;
    --T;
  }
  return T;
}

hipError_t launch_vcycle(hipStream_t st, BotState& b) {
  const int L = (int)b.mg.size() - 1;
  const int T = tail_first(b);  // levels T..L run in the one-block tail
  const int sep = T <= L ? T : L;                    // levels 0..sep-1 have launches of their own
  std::vector<double*> xf(b.mg.size());  // each level's iterate after its pre-smoothing, then after its post-smoothing
  std::vector<double*> xo(b.mg.size());  // the level's other buffer
  for (int l = 0; l < sep; ++l) {
    const BotState::MgLevel& lv = b.mg[l];
    const MgDev D = mgdev(lv);
    const double* r = l == 0 ? b.r : lv.r;
    // the sweeps' ping-pong between the level's two buffers: the way down leaves its iterate in y
    double *xa = lv.y, *xb = lv.x;
    const BotState::MgLevel& cv = b.mg[l + 1];
    const unsigned nbk = (unsigned)(((cv.a + CTX - 1) / CTX) * ((cv.b + CTY - 1) / CTY));
    hipLaunchKernelGGL(k_mg_down, dim3(nbk), dim3(BB), 0, st, D, mgdev(cv), r, xa, cv.r);
    BCHK(hipGetLastError());
    xf[l] = xa;
    xo[l] = xb;
  }
  if (T <= L) {
    MgTail t{};
    t.n = L - T + 1;
    size_t lds = 0;
    for (int q = 0; q < t.n; ++q) {
      const BotState::MgLevel& lv = b.mg[T + q];
      t.a[q] = lv.a;
      t.b[q] = lv.b;
      t.o[q] = lv.o;
      t.k[q] = lv.k;
      lds += 4 * sizeof(double) * (size_t)lv.a * lv.b;
    }
    t.dense_n = b.mg_n;
    t.inv = b.mg_inv;
    t.r0 = b.mg[T].r;
    // the buffer in which the per-level kernels leave level T's result (their ping-pong, replayed)
    double* x0 = b.mg[T].x;
    if (T < L) {
      double *xa = b.mg[T].x, *xb = b.mg[T].y;
      for (int s = 1; s < kNu; ++s) std::swap(xa, xb);
      for (int s = 1; s <= kNu; ++s) {
        double* out = xb;
        xb = xa;
        xa = out;
      }
      x0 = xa;
    }
    t.x0 = x0;
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailCells), lds, st, t);
    BCHK(hipGetLastError());
    xf[T] = x0;
  } else {
    const BotState::MgLevel& lv = b.mg[L];
    double* x = L == 0 ? b.z : lv.x;
    hipLaunchKernelGGL(k_mg_dense, dim3(1), dim3(BB), 0, st, mgdev(lv), b.mg_n, (const double*)b.mg_inv,
                       (const double*)(L == 0 ? b.r : lv.r), x);
    BCHK(hipGetLastError());
    xf[L] = x;
    if (L == 0) {
      const int n = b.nx * b.ny;
      hipLaunchKernelGGL(k_mg_dot, dim3(nblocks(n)), dim3(BB), 0, st, n, (const double*)b.r, (const double*)b.z,
                         b.part);
      return hipGetLastError();
    }
  }
  for (int l = sep - 1; l >= 0; --l) {
    const BotState::MgLevel& lv = b.mg[l];
    const MgDev D = mgdev(lv);
    const double* r = l == 0 ? b.r : lv.r;
    double* out = l == 0 ? b.z : xo[l];
    const unsigned nbk = (unsigned)(((lv.a + 2 * CTX - 1) / (2 * CTX)) * ((lv.b + 2 * CTY - 1) / (2 * CTY)));
    hipLaunchKernelGGL(k_mg_up, dim3(nbk), dim3(BB), 0, st, D, mgdev(b.mg[l + 1]), r, (const double*)xf[l],
                       (const double*)xf[l + 1], out);
    BCHK(hipGetLastError());
    xf[l] = out;
    if (l == 0) {
      hipLaunchKernelGGL(k_mg_dotm, dim3(nblocks((long long)lv.a * lv.b)), dim3(BB), 0, st, D, (const double*)b.r,
                         (const double*)b.z, b.part);
      BCHK(hipGetLastError());
    }
  }
  return hipSuccess;
}

hipError_t launch_chunk(hipStream_t st, BotState& b) {
  const BDims g = dims(b);
  const int n = b.nx * b.ny;
  for (int it = 0; it < kBotChunk; ++it) {
    double* dold = (it & 1) ? b.d1 : b.d0;
    double* dnew = (it & 1) ? b.d0 : b.d1;
    hipLaunchKernelGGL(k2_dirmv, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.S,
                       (const double*)b.z, (const double*)dold, dnew, b.q, b.part, (const double*)b.scal);
    BCHK(hipGetLastError());
    hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_ALPHA);
    BCHK(hipGetLastError());
    hipLaunchKernelGGL(k2_update, dim3(nblocks(n)), dim3(BB), 0, st, g, b.x, (const double*)dnew, b.r,
                       (const double*)b.q, (const double*)b.scal);
    BCHK(hipGetLastError());
    BCHK(launch_vcycle(st, b));
    hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_BETA);
    BCHK(hipGetLastError());
  }
  return hipSuccess;
}

// The multigrid hierarchy from the box operator K = -S A2 on the inner rows (host, once per handle):
// level 0 is the box with its edge excluded; a level's 2 x 2 aggregates of unknowns become the next
// level's cells, with K_c = P^T K P for the piecewise-constant P (5-point again: a coupling inside an
// aggregate adds to the coarse centre, one across its west/east/north/south side to that coarse
// coupling), until at most kMgDense unknowns remain; that operator is inverted (Gauss-Jordan with
// partial pivoting). Every loop runs in a fixed order, so the hierarchy is the same bits every time.
hipError_t mg_build(hipStream_t st, BotState& b) {
  const int nx = b.nx, ny = b.ny;
  const size_t n = (size_t)nx * ny;
  std::vector<double> coef(5 * n), S(n);
  BCHK(hipMemcpyAsync(coef.data(), b.coef, 5 * n * sizeof(double), hipMemcpyDeviceToHost, st));
  BCHK(hipMemcpyAsync(S.data(), b.S, n * sizeof(double), hipMemcpyDeviceToHost, st));
  BCHK(hipStreamSynchronize(st));
  struct HLevel {
    int a, b, o;
    std::vector<double> k;
  };
  std::vector<HLevel> lv;
  {
    HLevel L0{nx, ny, 1, std::vector<double>(5 * n, 0.0)};
    auto act = [&](int i, int j) { return i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1; };
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t c = (size_t)j * nx + i;
        if (!act(i, j)) {
          L0.k[c] = 1.0;
          continue;
        }
        L0.k[c] = -S[c] * coef[c];
        if (act(i - 1, j)) L0.k[n + c] = -S[c] * coef[n + c];
        if (act(i + 1, j)) L0.k[2 * n + c] = -S[c] * coef[2 * n + c];
        if (act(i, j - 1)) L0.k[3 * n + c] = -S[c] * coef[3 * n + c];
        if (act(i, j + 1)) L0.k[4 * n + c] = -S[c] * coef[4 * n + c];
      }
    // ||K - K^T||_F / ||K||_F over the inner rows: 0 when DLXP is constant along j and DLYP along i;
    // on other grids (spherical, curvilinear) K is not symmetric and CG is applied to it regardless
    double num = 0.0, den = 0.0;
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        const size_t c = (size_t)j * nx + i;
        for (int p = 0; p < 5; ++p) den += L0.k[p * n + c] * L0.k[p * n + c];
        // west against the west neighbour's east, north against the north neighbour's south
        if (act(i - 1, j)) {
          const double d = L0.k[n + c] - L0.k[2 * n + c - 1];
          num += 2.0 * d * d;
        }
        if (act(i, j - 1)) {
          const double d = L0.k[3 * n + c] - L0.k[4 * n + c - nx];
          num += 2.0 * d * d;
        }
      }
    b.sym_defect = den > 0.0 ? std::sqrt(num / den) : 0.0;
    lv.push_back(std::move(L0));
  }
  auto unknowns = [](const HLevel& L) { return (long long)(L.a - 2 * L.o) * (L.b - 2 * L.o); };
  while (unknowns(lv.back()) > kMgDense) {
    const HLevel& F = lv.back();
    const int wa = F.a - 2 * F.o, wb = F.b - 2 * F.o, ca = (wa + 1) / 2, cb = (wb + 1) / 2;
    const size_t fn = (size_t)F.a * F.b, cn = (size_t)ca * cb;
    HLevel C{ca, cb, 0, std::vector<double>(5 * cn, 0.0)};
    const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
    for (int j = F.o; j < F.b - F.o; ++j)
      for (int i = F.o; i < F.a - F.o; ++i) {
        const size_t f = (size_t)j * F.a + i;
        const int I = (i - F.o) / 2, J = (j - F.o) / 2;
        const size_t cc = (size_t)J * ca + I;
        C.k[cc] += F.k[f];
        for (int p = 1; p < 5; ++p) {
          const int ni = i + di[p], nj = j + dj[p];
          if (ni < F.o || nj < F.o || ni >= F.a - F.o || nj >= F.b - F.o) continue;
          const double v = F.k[p * fn + f];
          if ((ni - F.o) / 2 == I && (nj - F.o) / 2 == J)
            C.k[cc] += v;
          else
            C.k[p * cn + cc] += v;
        }
      }
    lv.push_back(std::move(C));
  }
  // the dense inverse of the last level over its unknowns, numbered row by row
  const HLevel& T = lv.back();
  const int wa = T.a - 2 * T.o, m = (int)unknowns(T);
  std::vector<double> M((size_t)m * m, 0.0), inv((size_t)m * m, 0.0);
  const size_t tn = (size_t)T.a * T.b;
  const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
  for (int t = 0; t < m; ++t) {
    const int i = t % wa + T.o, j = t / wa + T.o;
    const size_t c = (size_t)j * T.a + i;
    for (int p = 0; p < 5; ++p) {
      const int ni = i + di[p], nj = j + dj[p];
      if (ni < T.o || nj < T.o || ni >= T.a - T.o || nj >= T.b - T.o) continue;
      M[(size_t)t * m + (size_t)(nj - T.o) * wa + (ni - T.o)] += T.k[p * tn + c];
    }
    inv[(size_t)t * m + t] = 1.0;
  }
  for (int col = 0; col < m; ++col) {
    int piv = col;
    for (int rr = col + 1; rr < m; ++rr)
      if (std::fabs(M[(size_t)rr * m + col]) > std::fabs(M[(size_t)piv * m + col])) piv = rr;
    if (!(std::fabs(M[(size_t)piv * m + col]) > 0.0)) return hipErrorInvalidValue;
    if (piv != col)
      for (int q = 0; q < m; ++q) {
        std::swap(M[(size_t)piv * m + q], M[(size_t)col * m + q]);
        std::swap(inv[(size_t)piv * m + q], inv[(size_t)col * m + q]);
      }
    const double d = M[(size_t)col * m + col];
    for (int q = 0; q < m; ++q) {
      M[(size_t)col * m + q] /= d;
      inv[(size_t)col * m + q] /= d;
    }
    for (int rr = 0; rr < m; ++rr) {
      if (rr == col) continue;
      const double f = M[(size_t)rr * m + col];
      if (f == 0.0) continue;
      for (int q = 0; q < m; ++q) {
        M[(size_t)rr * m + q] -= f * M[(size_t)col * m + q];
        inv[(size_t)rr * m + q] -= f * inv[(size_t)col * m + q];
      }
    }
  }
  // device copies: level 0 uses the PCG's r and z plus two iterate buffers of the box
  auto dev = [&](double*& p, size_t cnt, const double* src) -> hipError_t {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, cnt * sizeof(double));
    if (e != hipSuccess) return e;
    b.mg_allocs.push_back(q);
    p = static_cast<double*>(q);
    if (src) return hipMemcpy(p, src, cnt * sizeof(double), hipMemcpyHostToDevice);
    return hipMemset(p, 0, cnt * sizeof(double));
  };
  b.mg.assign(lv.size(), BotState::MgLevel{});
  for (size_t l = 0; l < lv.size(); ++l) {
    BotState::MgLevel& d = b.mg[l];
    d.a = lv[l].a;
    d.b = lv[l].b;
    d.o = lv[l].o;
    const size_t cnt = (size_t)d.a * d.b;
    BCHK(dev(d.k, 5 * cnt, lv[l].k.data()));
    BCHK(dev(d.x, cnt, nullptr));
    BCHK(dev(d.y, cnt, nullptr));
    if (l > 0) BCHK(dev(d.r, cnt, nullptr));
  }
  BCHK(dev(b.mg_inv, (size_t)m * m, inv.data()));
  b.mg_n = m;
  return hipSuccess;
}

hipError_t read_scal(hipStream_t st, BotState& b) {
  BCHK(hipMemcpyAsync(b.scal_host, b.scal, S_N * sizeof(double), hipMemcpyDeviceToHost, st));
  return hipStreamSynchronize(st);
}

}  // namespace

static_assert(kBotChunk % 2 == 0, "the search directions alternate between two buffers");

hipError_t bot_setup(hipStream_t st, BotState& b) {
  const BDims g = dims(b);
  const int n = b.nx * b.ny;
  b.nblk = (int)nblocks(n);
  b.iters = -1;  // no solve yet: the first one starts from zero
  hipLaunchKernelGGL(k_bot_coef, dim3(nblocks(n)), dim3(BB), 0, st, g, b.dlxu, b.dlyv, b.dlxp, b.dlyp, b.coef, b.S);
  BCHK(hipGetLastError());
  BCHK(hipStreamSynchronize(st));
  if (b.nx > 2 && b.ny > 2) BCHK(mg_build(st, b));
  return hipSuccess;
}

hipError_t bot_step(hipStream_t st, BotState& b, const OcvArrays& a, const double* divg, double rtol, int maxit) {
  const BDims g = dims(b);
  const int n = b.nx * b.ny;
  const long long ncol = (long long)g.NU * g.JE + (long long)g.IE * g.NV + (long long)g.IE * g.JE;
  BCHK(hipEventRecord(b.ev0, st));
  hipLaunchKernelGGL(k_bot_extract, dim3(nblocks(ncol)), dim3(BB), 0, st, g, a, b.botgu, b.botgv, b.botgw);
  BCHK(hipGetLastError());
  hipLaunchKernelGGL(k_bot_rhs, dim3(nblocks(n)), dim3(BB), 0, st, g, a, (const double*)b.botgu,
                     (const double*)b.botgv, (const double*)b.botgw, divg, b.dbs, b.b);
  BCHK(hipGetLastError());
  // the first solve starts from zero, later ones from the previous solution (mo_petsc.f90:165); after a
  // solve that did not converge (a non-finite right-hand side or iterate among the causes) the next one
  // starts from zero again, so that one bad step does not poison every later one
  const int cold = (b.iters < 0 || b.status != ORCTM_NH_OK) ? 1 : 0;
  hipLaunchKernelGGL(k2_start, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.b,
                     b.x, b.d0, b.d1, cold);
  BCHK(hipGetLastError());
  hipLaunchKernelGGL(k2_resid, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.S,
                     (const double*)b.b, (const double*)b.x, b.r, b.part, 1);
  BCHK(hipGetLastError());
  hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_INIT);
  BCHK(hipGetLastError());
  const bool has_inner = b.nx > 2 && b.ny > 2;
  if (has_inner) {  // z = V-cycle(r) and r.z of the first iteration
    BCHK(launch_vcycle(st, b));
    hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_RZ);
    BCHK(hipGetLastError());
  }
  BCHK(read_scal(st, b));
  const double bb = b.scal_host[S_BB], tol2 = rtol * rtol * bb;
  double rr = b.scal_host[S_RR];
  int iters = 0, status = 0;
  if (!std::isfinite(bb) || !std::isfinite(rr)) status = ORCTM_NH_ERR_NONFINITE;
  while (status == 0 && has_inner && rr > tol2 && iters < maxit) {
    if (!b.chunk) {
      hipGraph_t graph = nullptr;
      BCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const hipError_t e1 = launch_chunk(st, b);
      const hipError_t e2 = hipStreamEndCapture(st, &graph);
      hipError_t e = e1 != hipSuccess ? e1 : e2;
      if (e == hipSuccess) e = hipGraphInstantiate(&b.chunk, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      BCHK(e);
    }
    BCHK(hipGraphLaunch(b.chunk, st));
    iters += kBotChunk;
    hipLaunchKernelGGL(k2_resid, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.S,
                       (const double*)b.b, (const double*)b.x, b.r, b.part, 0);
    BCHK(hipGetLastError());
    hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_CHECK);
    BCHK(hipGetLastError());
    BCHK(read_scal(st, b));
    rr = b.scal_host[S_RR];
    if (!std::isfinite(rr)) {
      status = ORCTM_NH_ERR_NONFINITE;
      break;
    }
    if (b.scal_host[S_FLAG] != 0.0 && !(rr <= tol2)) {
      status = ORCTM_NH_BREAKDOWN;
      break;
    }
  }
  if (status == 0 && !(rr <= tol2)) status = ORCTM_NH_NOT_CONVERGED;
  b.iters = iters;
  b.status = status;
  b.rel_res = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
  hipLaunchKernelGGL(k2_scatter, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.x, b.pnh_bot);
  BCHK(hipGetLastError());
  BCHK(launch_exchange(st, b.pnh_bot, b.IE, b.JE, 1, 1, 1, b.cyc_x, b.cyc_y));
  return hipEventRecord(b.ev1, st);
}

hipError_t bot_correct(hipStream_t st, const BotState& b, const OcvArrays& a, double* divg) {
  const BDims g = dims(b);
  hipLaunchKernelGGL(k_bot_correct, dim3(nblocks((long long)b.nx * b.ny)), dim3(BB), 0, st, g, b.dlxu, b.dlyv, b.dlxp,
                     b.dlyp, b.dzw, a.ddpo, a.weto, (const double*)b.pnh_bot, divg);
  return hipGetLastError();
}

void bot_release(BotState& b) {
  for (void* p : b.mg_allocs) hipFree(p);
  b.mg_allocs.clear();
  b.mg.clear();
  b.mg_inv = nullptr;
  if (b.chunk) hipGraphExecDestroy(b.chunk);
  b.chunk = nullptr;
  if (b.ev0) hipEventDestroy(b.ev0);
  if (b.ev1) hipEventDestroy(b.ev1);
  b.ev0 = b.ev1 = nullptr;
  if (b.scal_host) hipHostFree(b.scal_host);
  b.scal_host = nullptr;
}

}  // namespace onh
