// nh_bot.hip -- the 2-D bottom-pressure solve of the nonhydrostatic step, ORCTM's CG2D4NONHY option
// (ocvtot_nh.f90:64-147, mo_petsc.f90:147-208, 252-374, 401-538; off in the shipped Makefile), on the
// device, for one GPU holding the whole domain.
//
// Per step, after DIVG:
//   k_bot_extract   BOTGU, BOTGV, BOTGW: the tendency at the bottom wet level of every column, as the
//                   source's K loop assigns it (one thread per column walks K)
//   k_bot_rhs       DIVG_bot_surf = bottom divergence - DIVG(:,:,1)*WETO on the interior
//   2-D solve       compute_matrix_2d's operator on the (IE-2) x (JE-2) interior box: the box-edge
//                   rows (diagonal only, Neumann = .FALSE.) directly, then Jacobi-preconditioned CG on
//                   the remaining rows scaled by S = DLXP*DLYP, warm-started from the previous
//                   solution as KSPSetInitialGuessNonzero does (mo_petsc.f90:165), to
//                   ||b - A2 p|| <= rtol ||b|| checked every kBotChunk iterations; block sums in fixed
//                   order, so the iterates are the same bits on every run
//   k2_scatter      PNH_bot on the interior, then bounds_exch(PNH_bot) (ocvtot_nh.f90:146)
//   k_bot_correct   ComputeRHS's corrections of DIVG by PNH_bot at dry neighbours and the bottom face
// FP64 throughout; the bottom values, DIVG_bot_surf and the corrections evaluate each expression left
// to right as the source writes it (no contraction), so they are the oracle's bits.
#include <hip/hip_runtime.h>

#include <cmath>

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

// t = b - A2 x on every row; partial sums of t^2 and b^2; with init also r = -S t, z = r / (-S A_cc) on the
// inner rows and the partial sum of r z.
__global__ __launch_bounds__(BB) void k2_resid(BDims g, const double* __restrict__ coef,
                                               const double* __restrict__ S, const double* __restrict__ b,
                                               const double* __restrict__ x, double* __restrict__ r,
                                               double* __restrict__ z, double* __restrict__ part, int init) {
  const int c = blockIdx.x * BB + threadIdx.x;
  const int n = g.nx * g.ny;
  double tt = 0.0, bb = 0.0, rz = 0.0;
  if (c < n) {
    double ax;
    const bool in = g.inner(c);
    if (in) ax = apply_inner(g, coef, c, x[c - g.nx], x[c - 1], x[c], x[c + 1], x[c + g.nx]);
    else ax = coef[c] * x[c];
    const double t = b[c] - ax;
    tt = t * t;
    bb = b[c] * b[c];
    if (init) {
      const double rc = in ? -S[c] * t : 0.0;
      const double zc = in ? rc / (-S[c] * coef[c]) : 0.0;
      r[c] = rc;
      z[c] = zc;
      rz = rc * zc;
    }
  }
  block_sums(tt, bb, rz, part);
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

// x += alpha d, r -= alpha q, z = r / (-S A_cc) on the inner rows; partial sum of r z.
__global__ __launch_bounds__(BB) void k2_update(BDims g, const double* __restrict__ coef,
                                                const double* __restrict__ S, double* __restrict__ x,
                                                const double* __restrict__ d, double* __restrict__ r,
                                                const double* __restrict__ q, double* __restrict__ z,
                                                double* __restrict__ part, const double* __restrict__ scal) {
  const int c = blockIdx.x * BB + threadIdx.x;
  const double alpha = scal[S_ALPHA];
  double rz = 0.0;
  if (c < g.nx * g.ny && g.inner(c)) {
    x[c] += alpha * d[c];
    const double rc = r[c] - alpha * q[c];
    const double zc = rc / (-S[c] * coef[c]);
    r[c] = rc;
    z[c] = zc;
    rz = rc * zc;
  }
  block_sums(rz, 0.0, 0.0, part);
}

enum { R_INIT = 0, R_ALPHA, R_BETA, R_CHECK };

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
    hipLaunchKernelGGL(k2_update, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.S,
                       b.x, (const double*)dnew, b.r, (const double*)b.q, b.z, b.part, (const double*)b.scal);
    BCHK(hipGetLastError());
    hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_BETA);
    BCHK(hipGetLastError());
  }
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
  return hipStreamSynchronize(st);
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
  // the first solve starts from zero, later ones from the previous solution (mo_petsc.f90:165)
  const int cold = b.iters < 0 ? 1 : 0;
  hipLaunchKernelGGL(k2_start, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.b,
                     b.x, b.d0, b.d1, cold);
  BCHK(hipGetLastError());
  hipLaunchKernelGGL(k2_resid, dim3(nblocks(n)), dim3(BB), 0, st, g, (const double*)b.coef, (const double*)b.S,
                     (const double*)b.b, (const double*)b.x, b.r, b.z, b.part, 1);
  BCHK(hipGetLastError());
  hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_INIT);
  BCHK(hipGetLastError());
  BCHK(read_scal(st, b));
  const double bb = b.scal_host[S_BB], tol2 = rtol * rtol * bb;
  double rr = b.scal_host[S_RR];
  int iters = 0, status = 0;
  const bool has_inner = b.nx > 2 && b.ny > 2;
  while (has_inner && rr > tol2 && iters < maxit) {
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
                       (const double*)b.b, (const double*)b.x, b.r, b.z, b.part, 0);
    BCHK(hipGetLastError());
    hipLaunchKernelGGL(k2_reduce, dim3(1), dim3(BB), 0, st, (const double*)b.part, b.nblk, b.scal, (int)R_CHECK);
    BCHK(hipGetLastError());
    BCHK(read_scal(st, b));
    rr = b.scal_host[S_RR];
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
  if (b.chunk) hipGraphExecDestroy(b.chunk);
  b.chunk = nullptr;
  if (b.ev0) hipEventDestroy(b.ev0);
  if (b.ev1) hipEventDestroy(b.ev1);
  b.ev0 = b.ev1 = nullptr;
  if (b.scal_host) hipHostFree(b.scal_host);
  b.scal_host = nullptr;
}

}  // namespace onh
