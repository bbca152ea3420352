/*
 * nh_oracle_bot.c -- CPU restatement of ORCTM's 2-D bottom-pressure solve, the CG2D4NONHY option of
 * the nonhydrostatic step (ORCTM/src_orctm/mo_petsc.f90, ocvtot_nh.f90; off in the shipped Makefile,
 * src_orctm/Makefile:48 -DZZCG2D4NONHY). TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's
 * CPU-baseline leg, never by the product. One process holds the whole domain (have_g_* all true).
 *
 *   orc_bottom_rhs         ocvtot_nh.f90:64-144   bottom values of GU, GV, GW, their divergence,
 *                                                 DIVG_bot_surf
 *   orc_op2d_coef/_apply   mo_petsc.f90:401-538   compute_matrix_2d with Neumann = .FALSE. (:37)
 *   orc_solve2d            the device's PCG (no counterpart in the reference: its cg2d_ KSP runs
 *                          inside PETSc, mo_petsc.f90:147-166)
 *   orc_bottom_correction  mo_petsc.f90:252-374   ComputeRHS's corrections of DIVG by PNH_bot
 *
 * Every expression is evaluated left to right as the Fortran source writes it (no contraction).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "nh_oracle_bot.h"

#define P2(i, j) ((size_t)((i)-1) + (size_t)IE * (size_t)((j)-1))
#define P3(i, j, k) ((size_t)((i)-1) + (size_t)IE * ((size_t)((j)-1) + (size_t)JE * (size_t)((k)-1)))
#define U2(i, j) ((size_t)((i)-IS) + (size_t)NU * (size_t)((j)-1))
#define U3(i, j, k) ((size_t)((i)-IS) + (size_t)NU * ((size_t)((j)-1) + (size_t)JE * (size_t)((k)-1)))
#define V2(i, j) ((size_t)((i)-1) + (size_t)IE * (size_t)((j)-JS))
#define V3(i, j, k) ((size_t)((i)-1) + (size_t)IE * ((size_t)((j)-JS) + (size_t)NV * (size_t)((k)-1)))

/*
 * ocvtot_nh.f90:64-144. For each (I,J) the K loop assigns, in order: K = 1 and the top face dry: 0;
 * K = KE and the bottom face wet: G(KE); otherwise at a wet-to-dry step between K-1 and K: G(K-1). The
 * last assignment stands; with none the previous value stays. BOTGW divides by DDPO of that level.
 * DIVBOTGU = (BOTGU(I)-BOTGU(I-1))/DLXP, DIVBOTGV = (BOTGV(J-1)-BOTGV(J))/DLYP, DIVBOTGW = BOTGW,
 * DIVG_bot_surf = (DIVBOTGU + DIVBOTGV + DIVBOTGW) - DIVG(I,J,1)*WETO(I,J,1).
 */
void orc_bottom_rhs(const orc_vgrid* g, const double* uoo, const double* uoo_1, const double* voe,
                    const double* voe_1, const double* wno, const double* wno_1, const double* divg, double* botgu,
                    double* botgv, double* botgw, double* dbs) {
  const int IE = g->IE, JE = g->JE, KE = g->KE, IS = g->i_start, JS = g->j_start;
  const int NU = IE - IS + 1, NV = JE - JS + 1;
  const double dt = g->dt;
  for (int K = 1; K <= KE; ++K)
    for (int J = 1; J <= JE; ++J)
      for (int I = IS; I <= IE; ++I) {
        const double am = g->amsuo[U3(I, J, K)];
        if (K == 1) {
          if (am == 0.0) botgu[U2(I, J)] = 0.0;
        } else if (K == KE) {
          if (am != 0.0) botgu[U2(I, J)] = (uoo[U3(I, J, K)] - uoo_1[U3(I, J, K)]) / dt;
        } else if (g->amsuo[U3(I, J, K - 1)] != 0.0 && am == 0.0) {
          botgu[U2(I, J)] = (uoo[U3(I, J, K - 1)] - uoo_1[U3(I, J, K - 1)]) / dt;
        }
      }
  for (int K = 1; K <= KE; ++K)
    for (int J = JS; J <= JE; ++J)
      for (int I = 1; I <= IE; ++I) {
        const double am = g->amsue[V3(I, J, K)];
        if (K == 1) {
          if (am == 0.0) botgv[V2(I, J)] = 0.0;
        } else if (K == KE) {
          if (am != 0.0) botgv[V2(I, J)] = (voe[V3(I, J, K)] - voe_1[V3(I, J, K)]) / dt;
        } else if (g->amsue[V3(I, J, K - 1)] != 0.0 && am == 0.0) {
          botgv[V2(I, J)] = (voe[V3(I, J, K - 1)] - voe_1[V3(I, J, K - 1)]) / dt;
        }
      }
  for (int K = 1; K <= KE; ++K)
    for (int J = 1; J <= JE; ++J)
      for (int I = 1; I <= IE; ++I) {
        const double w = g->weto[P3(I, J, K)];
        if (K == 1) {
          if (w == 0.0) botgw[P2(I, J)] = 0.0;
        } else if (K == KE) {
          if (w != 0.0) botgw[P2(I, J)] = (wno[P3(I, J, K)] - wno_1[P3(I, J, K)]) / dt / g->ddpo[P3(I, J, K)];
        } else if (g->weto[P3(I, J, K - 1)] != 0.0 && w == 0.0) {
          botgw[P2(I, J)] = (wno[P3(I, J, K - 1)] - wno_1[P3(I, J, K - 1)]) / dt / g->ddpo[P3(I, J, K - 1)];
        }
      }
  for (int J = 2; J <= JE - 1; ++J)
    for (int I = 2; I <= IE - 1; ++I) {
      const double du = (botgu[U2(I, J)] - botgu[U2(I - 1, J)]) / g->dlxp[P2(I, J)];
      const double dv = (botgv[V2(I, J - 1)] - botgv[V2(I, J)]) / g->dlyp[P2(I, J)];
      const double dw = botgw[P2(I, J)];
      dbs[P2(I, J)] = (du + dv + dw) - divg[P3(I, J, 1)] * g->weto[P3(I, J, 1)];
    }
}

/* mo_petsc.f90:415-527. XU = 1/DLXU, XP = 1/DLXP, YV = 1/DLYV, YP = 1/DLYP (:414-421);
 * X_W = XU(I-1)XP, X_C = (XU(I-1)+XU(I))XP, X_E = XU(I)XP, Y_N = YV(J-1)YP, Y_C = (YV(J-1)+YV(J))YP,
 * Y_S = YV(J)YP; a row on the edge of the box (DMDA index 0 or m-1 in x or y) gets -(X_C+Y_C) alone
 * (:485-487), every other row W, N, -(X_C+Y_C), S, E (:490-518). */
void orc_op2d_coef(const orctm_nh_grid_t* g, double* coef) {
  const int IE = g->ie, JE = g->je, IS = g->i_start, JS = g->j_start, NU = IE - IS + 1;
  const int nx = IE - 2, ny = JE - 2;
  const size_t n = (size_t)nx * ny;
  (void)JS;
  for (int J = 2; J <= JE - 1; ++J)
    for (int I = 2; I <= IE - 1; ++I) {
      const double xuw = 1.0 / g->dlxu[U2(I - 1, J)], xue = 1.0 / g->dlxu[U2(I, J)], xp = 1.0 / g->dlxp[P2(I, J)];
      const double yvn = 1.0 / g->dlyv[V2(I, J - 1)], yvs = 1.0 / g->dlyv[V2(I, J)], yp = 1.0 / g->dlyp[P2(I, J)];
      const double X_W = xuw * xp, X_C = (xuw + xue) * xp, X_E = xue * xp;
      const double Y_N = yvn * yp, Y_C = (yvn + yvs) * yp, Y_S = yvs * yp;
      const size_t c = (size_t)(I - 2) + (size_t)nx * (J - 2);
      const int edge = I == 2 || J == 2 || I == IE - 1 || J == JE - 1;
      coef[c] = -(X_C + Y_C);
      coef[n + c] = edge ? 0.0 : X_W;
      coef[2 * n + c] = edge ? 0.0 : X_E;
      coef[3 * n + c] = edge ? 0.0 : Y_N;
      coef[4 * n + c] = edge ? 0.0 : Y_S;
    }
}

void orc_op2d_apply(const orctm_nh_grid_t* g, const double* coef, const double* p, double* y) {
  const int nx = g->ie - 2, ny = g->je - 2;
  const size_t n = (size_t)nx * ny;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t c = (size_t)i + (size_t)nx * j;
      double s = 0.0;
      if (j > 0) s += coef[3 * n + c] * p[c - nx];
      if (i > 0) s += coef[n + c] * p[c - 1];
      s += coef[c] * p[c];
      if (i < nx - 1) s += coef[2 * n + c] * p[c + 1];
      if (j < ny - 1) s += coef[4 * n + c] * p[c + nx];
      y[c] = s;
    }
}

static double norm2(const double* v, size_t n) {
  double s = 0.0;
  for (size_t c = 0; c < n; ++c) s += v[c] * v[c];
  return sqrt(s);
}

int orc_solve2d(const orctm_nh_grid_t* g, const double* coef, const double* b, double* x, double rtol, int maxit,
                double* rel_res) {
  const int IE = g->ie, JE = g->je, nx = IE - 2, ny = JE - 2;
  const size_t n = (size_t)nx * ny;
  double* w = (double*)calloc(7 * n, sizeof(double));
  double *r = w, *z = w + n, *d = w + 2 * n, *q = w + 3 * n, *S = w + 4 * n, *ax = w + 5 * n, *inner = w + 6 * n;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t c = (size_t)i + (size_t)nx * j;
      const int I = i + 2, J = j + 2;
      inner[c] = (i > 0 && j > 0 && i < nx - 1 && j < ny - 1) ? 1.0 : 0.0;
      S[c] = g->dlxp[P2(I, J)] * g->dlyp[P2(I, J)];
      if (inner[c] == 0.0) x[c] = b[c] / coef[c];
    }
  const double bn = norm2(b, n);
  orc_op2d_apply(g, coef, x, ax);
  for (size_t c = 0; c < n; ++c) q[c] = b[c] - ax[c];
  double rr = norm2(q, n);
  int it = 0;
  if (rr > rtol * bn && nx > 2 && ny > 2) {
    /* PCG on the inner rows of C = -S*A (symmetric when DLYP is constant along i and DLXP along j),
       Jacobi preconditioner diag(C) = -S*A_cc; the box-edge values enter through the residual */
    double rz = 0.0;
    for (size_t c = 0; c < n; ++c) {
      r[c] = inner[c] != 0.0 ? -S[c] * q[c] : 0.0;
      z[c] = inner[c] != 0.0 ? r[c] / (-S[c] * coef[c]) : 0.0;
      d[c] = z[c];
      rz += r[c] * z[c];
    }
    while (it < maxit) {
      orc_op2d_apply(g, coef, d, ax);
      double dq = 0.0;
      for (size_t c = 0; c < n; ++c) {
        q[c] = inner[c] != 0.0 ? -S[c] * ax[c] : 0.0;
        dq += d[c] * q[c];
      }
      if (!(dq > 0.0)) break;
      const double alpha = rz / dq;
      double rzn = 0.0;
      for (size_t c = 0; c < n; ++c) {
        x[c] += alpha * d[c];
        r[c] -= alpha * q[c];
        z[c] = inner[c] != 0.0 ? r[c] / (-S[c] * coef[c]) : 0.0;
        rzn += r[c] * z[c];
      }
      ++it;
      if (it % 32 == 0 || it == maxit) {
        orc_op2d_apply(g, coef, x, ax);
        for (size_t c = 0; c < n; ++c) q[c] = b[c] - ax[c];
        rr = norm2(q, n);
        if (rr <= rtol * bn) break;
      }
      const double beta = rzn / rz;
      rz = rzn;
      for (size_t c = 0; c < n; ++c) d[c] = z[c] + beta * d[c];
    }
    orc_op2d_apply(g, coef, x, ax);
    for (size_t c = 0; c < n; ++c) q[c] = b[c] - ax[c];
    rr = norm2(q, n);
  }
  if (rel_res) *rel_res = bn > 0.0 ? rr / bn : rr;
  free(w);
  return it;
}

/*
 * mo_petsc.f90:252-374 with CG2D4NONHY. Pre-pass per column (:252-292): ZW(K) = 1/dddwo with dddwo
 * DDPO(1)/2 (DZW(1)/2 if DDPO(1) = 0) at K = 1, DZW(KE)-DDPO(KE)/2 (DZW(KE)/2 if DDPO(KE) = 0) at
 * K = KEP, and as compute_matrix between; ZP(K) = 1/((1/ZW(K) + 1/ZW(K+1))/2); XU, XP, YV, YP as in
 * compute_matrix_2d. Per cell (:302-374): a dry cell adds (X_C+Y_C+Z_C)*PNH_bot(I,J); a wet cell
 * subtracts Z_U*PNH_bot(I,J) (Z_U = 0 at K = 1) under a dry cell above, Y_N*PNH_bot(I,J-1),
 * X_W*PNH_bot(I-1,J), X_E*PNH_bot(I+1,J), Y_S*PNH_bot(I,J+1) for dry neighbours, and Z_D*PNH_bot(I,J)
 * at K = KE or above a dry cell, in that order.
 */
void orc_bottom_correction(const orctm_nh_grid_t* g, const double* pnh_bot, double* divg) {
  const int IE = g->ie, JE = g->je, KE = g->ke, KEP = KE + 1, IS = g->i_start, JS = g->j_start;
  const int NU = IE - IS + 1;
  (void)JS;
  double* zw = (double*)malloc(sizeof(double) * (size_t)(2 * KEP));
  double* zp = zw + KEP;
  for (int J = 2; J <= JE - 1; ++J)
    for (int I = 2; I <= IE - 1; ++I) {
      for (int K = 1; K <= KEP; ++K) {
        double dd;
        if (K == 1) {
          dd = (g->ddpo[P3(I, J, 1)] == 0.0) ? g->dzw[0] / 2.0 : g->ddpo[P3(I, J, 1)] / 2.0;
        } else if (K == KEP) {
          dd = (g->ddpo[P3(I, J, KE)] == 0.0) ? g->dzw[KE - 1] / 2.0 : g->dzw[KE - 1] - g->ddpo[P3(I, J, KE)] / 2.0;
        } else {
          const double dm = g->ddpo[P3(I, J, K - 1)], d0 = g->ddpo[P3(I, J, K)];
          if (dm == 0.0)
            dd = (g->dzw[K - 1] + g->dzw[K - 2]) / 2.0;
          else if (d0 == 0.0)
            dd = dm / 2.0;
          else
            dd = (dm + d0) / 2.0;
        }
        zw[K - 1] = 1.0 / dd;
      }
      for (int K = 1; K <= KE; ++K) zp[K - 1] = 1.0 / ((1.0 / zw[K - 1] + 1.0 / zw[K]) / 2.0);
      const double xuw = 1.0 / g->dlxu[U2(I - 1, J)], xue = 1.0 / g->dlxu[U2(I, J)], xp = 1.0 / g->dlxp[P2(I, J)];
      const double yvn = 1.0 / g->dlyv[V2(I, J - 1)], yvs = 1.0 / g->dlyv[V2(I, J)], yp = 1.0 / g->dlyp[P2(I, J)];
      const double X_W = xuw * xp, X_C = (xuw + xue) * xp, X_E = xue * xp;
      const double Y_N = yvn * yp, Y_C = (yvn + yvs) * yp, Y_S = yvs * yp;
      const double pc = pnh_bot[P2(I, J)];
      for (int K = 1; K <= KE; ++K) {
        const double Z_C = (zw[K - 1] + zw[K]) * zp[K - 1], Z_D = zw[K] * zp[K - 1];
        double Z_U = zw[K - 1] * zp[K - 1];
        double* dv = &divg[P3(I, J, K)];
        if (g->weto[P3(I, J, K)] == 0.0) {
          *dv = *dv + (X_C + Y_C + Z_C) * pc;
          continue;
        }
        if (K == 1) Z_U = 0.0;
        if (K == 1)
          *dv = *dv - Z_U * pc;
        else if (g->weto[P3(I, J, K - 1)] == 0.0)
          *dv = *dv - Z_U * pc;
        if (g->weto[P3(I, J - 1, K)] == 0.0) *dv = *dv - Y_N * pnh_bot[P2(I, J - 1)];
        if (g->weto[P3(I - 1, J, K)] == 0.0) *dv = *dv - X_W * pnh_bot[P2(I - 1, J)];
        if (g->weto[P3(I + 1, J, K)] == 0.0) *dv = *dv - X_E * pnh_bot[P2(I + 1, J)];
        if (g->weto[P3(I, J + 1, K)] == 0.0) *dv = *dv - Y_S * pnh_bot[P2(I, J + 1)];
        if (K == KE)
          *dv = *dv - Z_D * pc;
        else if (g->weto[P3(I, J, K + 1)] == 0.0)
          *dv = *dv - Z_D * pc;
      }
    }
  free(zw);
}
