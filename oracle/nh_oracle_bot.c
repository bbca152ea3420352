/*
 * nh_oracle_bot.c -- CPU restatement of ORCTM's 2-D bottom-pressure solve, the CG2D4NONHY option of
 * the nonhydrostatic step (ORCTM/src_orctm/mo_petsc.f90, ocvtot_nh.f90; off in the shipped Makefile,
 * src_orctm/Makefile:48 -DZZCG2D4NONHY). TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's
 * CPU-baseline leg, never by the product. One process holds the whole domain (have_g_* all true).
 *
 *   orc_bottom_rhs         ocvtot_nh.f90:64-144   bottom values of GU, GV, GW, their divergence,
 *                                                 DIVG_bot_surf
 *   orc_op2d_coef/_apply   mo_petsc.f90:401-538   compute_matrix_2d with Neumann = .FALSE. (:37)
 *   orc_solve2d            the device's PCG with its multigrid V-cycle (no counterpart in the
 *                          reference: its cg2d_ KSP runs inside PETSc, mo_petsc.f90:147-166)
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

/* The multigrid preconditioner: level 0 is the box (unknowns i in [1, nx-2], j in [1, ny-2]) with
 * K = -S*A2 on them (couplings to the box edge dropped); each next level's cell is a 2 x 2 aggregate of
 * the unknowns before it, with the Galerkin operator P^T K P of the piecewise-constant P (a coupling
 * inside an aggregate adds to the coarse centre, one across a side to that side's coarse coupling);
 * the last level (at most MG_DENSE unknowns) is inverted by Gauss-Jordan with partial pivoting. */
#define MG_DENSE 64
#define MG_OMEGA (2.0 / 3.0)
#define MG_ALPHA 1.8
#define MG_NU 2
#define MG_CHECK 8

typedef struct {
  int a, b, o;
  double *k, *x, *y, *r;
} mglev;

static int mg_active(const mglev* L, int i, int j) { return i >= L->o && j >= L->o && i < L->a - L->o && j < L->b - L->o; }

static double mg_kx(const mglev* L, const double* x, int i, int j, size_t c) {
  const size_t n = (size_t)L->a * L->b;
  double s = L->k[c] * x[c];
  if (i > 0) s += L->k[n + c] * x[c - 1];
  if (i < L->a - 1) s += L->k[2 * n + c] * x[c + 1];
  if (j > 0) s += L->k[3 * n + c] * x[c - L->a];
  if (j < L->b - 1) s += L->k[4 * n + c] * x[c + L->a];
  return s;
}

static long mg_unknowns(const mglev* L) { return (long)(L->a - 2 * L->o) * (L->b - 2 * L->o); }

/* y = x + omega (r - K x) / K_cc on the unknowns (x = 0 when x is NULL), 0 elsewhere */
static void mg_jac(const mglev* L, const double* r, const double* x, double* y) {
  for (int j = 0; j < L->b; ++j)
    for (int i = 0; i < L->a; ++i) {
      const size_t c = (size_t)j * L->a + i;
      if (!mg_active(L, i, j))
        y[c] = 0.0;
      else if (!x)
        y[c] = MG_OMEGA * r[c] / L->k[c];
      else
        y[c] = x[c] + MG_OMEGA * (r[c] - mg_kx(L, x, i, j, c)) / L->k[c];
    }
}

typedef struct {
  int nl, m;
  mglev* lv;
  double* inv;
} mghier;

static void mg_free(mghier* h) {
  for (int l = 0; l < h->nl; ++l) {
    free(h->lv[l].k);
    free(h->lv[l].x);
    free(h->lv[l].y);
    free(h->lv[l].r);
  }
  free(h->lv);
  free(h->inv);
}

static int mg_build(int nx, int ny, const double* coef, const double* S, mghier* h) {
  const size_t n = (size_t)nx * ny;
  const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
  h->nl = 1;
  h->lv = (mglev*)calloc(64, sizeof(mglev));
  mglev* L0 = &h->lv[0];
  L0->a = nx;
  L0->b = ny;
  L0->o = 1;
  L0->k = (double*)calloc(5 * n, sizeof(double));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t c = (size_t)j * nx + i;
      if (!mg_active(L0, i, j)) {
        L0->k[c] = 1.0;
        continue;
      }
      L0->k[c] = -S[c] * coef[c];
      for (int p = 1; p < 5; ++p)
        if (mg_active(L0, i + di[p], j + dj[p])) L0->k[p * n + c] = -S[c] * coef[p * n + c];
    }
  while (mg_unknowns(&h->lv[h->nl - 1]) > MG_DENSE) {
    const mglev* F = &h->lv[h->nl - 1];
    mglev* C = &h->lv[h->nl];
    const int wa = F->a - 2 * F->o, wb = F->b - 2 * F->o;
    C->a = (wa + 1) / 2;
    C->b = (wb + 1) / 2;
    C->o = 0;
    const size_t fn = (size_t)F->a * F->b, cn = (size_t)C->a * C->b;
    C->k = (double*)calloc(5 * cn, sizeof(double));
    for (int j = F->o; j < F->b - F->o; ++j)
      for (int i = F->o; i < F->a - F->o; ++i) {
        const size_t f = (size_t)j * F->a + i;
        const int I = (i - F->o) / 2, J = (j - F->o) / 2;
        const size_t cc = (size_t)J * C->a + I;
        C->k[cc] += F->k[f];
        for (int p = 1; p < 5; ++p) {
          const int ni = i + di[p], nj = j + dj[p];
          if (!mg_active(F, ni, nj)) continue;
          if ((ni - F->o) / 2 == I && (nj - F->o) / 2 == J)
            C->k[cc] += F->k[p * fn + f];
          else
            C->k[p * cn + cc] += F->k[p * fn + f];
        }
      }
    h->nl++;
  }
  for (int l = 0; l < h->nl; ++l) {
    const size_t cnt = (size_t)h->lv[l].a * h->lv[l].b;
    h->lv[l].x = (double*)calloc(cnt, sizeof(double));
    h->lv[l].y = (double*)calloc(cnt, sizeof(double));
    h->lv[l].r = (double*)calloc(cnt, sizeof(double));
  }
  const mglev* T = &h->lv[h->nl - 1];
  const int wa = T->a - 2 * T->o, m = (int)mg_unknowns(T);
  const size_t tn = (size_t)T->a * T->b;
  double* M = (double*)calloc((size_t)m * m, sizeof(double));
  h->inv = (double*)calloc((size_t)m * m, sizeof(double));
  h->m = m;
  for (int t = 0; t < m; ++t) {
    const int i = t % wa + T->o, j = t / wa + T->o;
    const size_t c = (size_t)j * T->a + i;
    for (int p = 0; p < 5; ++p) {
      const int ni = i + di[p], nj = j + dj[p];
      if (!mg_active(T, ni, nj)) continue;
      M[(size_t)t * m + (size_t)(nj - T->o) * wa + (ni - T->o)] += T->k[p * tn + c];
    }
    h->inv[(size_t)t * m + t] = 1.0;
  }
  int ok = 1;
  for (int col = 0; col < m && ok; ++col) {
    int piv = col;
    for (int rr = col + 1; rr < m; ++rr)
      if (fabs(M[(size_t)rr * m + col]) > fabs(M[(size_t)piv * m + col])) piv = rr;
    if (!(fabs(M[(size_t)piv * m + col]) > 0.0)) {
      ok = 0;
      break;
    }
    if (piv != col)
      for (int q = 0; q < m; ++q) {
        double t1 = M[(size_t)piv * m + q];
        M[(size_t)piv * m + q] = M[(size_t)col * m + q];
        M[(size_t)col * m + q] = t1;
        t1 = h->inv[(size_t)piv * m + q];
        h->inv[(size_t)piv * m + q] = h->inv[(size_t)col * m + q];
        h->inv[(size_t)col * m + q] = t1;
      }
    const double d = M[(size_t)col * m + col];
    for (int q = 0; q < m; ++q) {
      M[(size_t)col * m + q] /= d;
      h->inv[(size_t)col * m + q] /= d;
    }
    for (int rr = 0; rr < m; ++rr) {
      if (rr == col) continue;
      const double f = M[(size_t)rr * m + col];
      if (f == 0.0) continue;
      for (int q = 0; q < m; ++q) {
        M[(size_t)rr * m + q] -= f * M[(size_t)col * m + q];
        h->inv[(size_t)rr * m + q] -= f * h->inv[(size_t)col * m + q];
      }
    }
  }
  free(M);
  return ok;
}

/* z = V-cycle(r) on level l: kNu damped-Jacobi sweeps, the residual summed over each aggregate, the
 * coarse solution scaled by MG_ALPHA added back, kNu sweeps; the dense inverse on the last level */
static void mg_vcycle(mghier* h, int l, const double* r, double* z) {
  mglev* L = &h->lv[l];
  const size_t cnt = (size_t)L->a * L->b;
  if (l == h->nl - 1) {
    const int wa = L->a - 2 * L->o, m = h->m;
    for (int j = 0; j < L->b; ++j)
      for (int i = 0; i < L->a; ++i) {
        const size_t c = (size_t)j * L->a + i;
        double s = 0.0;
        if (mg_active(L, i, j)) {
          const int t = (j - L->o) * wa + (i - L->o);
          for (int q = 0; q < m; ++q) s += h->inv[(size_t)t * m + q] * r[(size_t)(q / wa + L->o) * L->a + (q % wa + L->o)];
        }
        z[c] = s;
      }
    return;
  }
  double *xa = L->x, *xb = L->y;
  mg_jac(L, r, NULL, xa);
  for (int s = 1; s < MG_NU; ++s) {
    mg_jac(L, r, xa, xb);
    double* t = xa;
    xa = xb;
    xb = t;
  }
  mglev* C = &h->lv[l + 1];
  for (int J = 0; J < C->b; ++J)
    for (int I = 0; I < C->a; ++I) {
      double s = 0.0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int i = L->o + 2 * I + di, j = L->o + 2 * J + dj;
          if (mg_active(L, i, j)) {
            const size_t f = (size_t)j * L->a + i;
            s += r[f] - mg_kx(L, xa, i, j, f);
          }
        }
      C->r[(size_t)J * C->a + I] = s;
    }
  double* zc = (double*)calloc((size_t)C->a * C->b, sizeof(double));
  mg_vcycle(h, l + 1, C->r, zc);
  for (int j = L->o; j < L->b - L->o; ++j)
    for (int i = L->o; i < L->a - L->o; ++i) xa[(size_t)j * L->a + i] += MG_ALPHA * zc[(size_t)((j - L->o) / 2) * C->a + (i - L->o) / 2];
  free(zc);
  for (int s = 1; s <= MG_NU; ++s) {
    double* out = (s == MG_NU) ? z : xb;
    mg_jac(L, r, xa, out);
    xb = xa;
    xa = out;
  }
  (void)cnt;
}

int orc_mg_apply(const orctm_nh_grid_t* g, const double* coef, const double* r, double* z) {
  const int IE = g->ie, JE = g->je, nx = IE - 2, ny = JE - 2;
  const size_t n = (size_t)nx * ny;
  if (nx <= 2 || ny <= 2) return 0;
  double* S = (double*)calloc(n, sizeof(double));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) S[(size_t)i + (size_t)nx * j] = g->dlxp[P2(i + 2, j + 2)] * g->dlyp[P2(i + 2, j + 2)];
  mghier h;
  memset(&h, 0, sizeof(h));
  const int ok = mg_build(nx, ny, coef, S, &h);
  if (ok) mg_vcycle(&h, 0, r, z);
  const int nl = h.nl;
  mg_free(&h);
  free(S);
  return ok ? nl : 0;
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
  mghier h;
  memset(&h, 0, sizeof(h));
  if (rr > rtol * bn && nx > 2 && ny > 2 && mg_build(nx, ny, coef, S, &h)) {
    /* PCG on the inner rows of C = -S*A (symmetric when DLYP is constant along i and DLXP along j),
       preconditioned by one V-cycle; the box-edge values enter through the residual */
    for (size_t c = 0; c < n; ++c) r[c] = inner[c] != 0.0 ? -S[c] * q[c] : 0.0;
    mg_vcycle(&h, 0, r, z);
    double rz = 0.0;
    for (size_t c = 0; c < n; ++c) {
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
      for (size_t c = 0; c < n; ++c) {
        if (inner[c] == 0.0) continue;
        x[c] += alpha * d[c];
        r[c] -= alpha * q[c];
      }
      mg_vcycle(&h, 0, r, z);
      double rzn = 0.0;
      for (size_t c = 0; c < n; ++c) rzn += r[c] * z[c];
      ++it;
      if (it % MG_CHECK == 0 || it == maxit) {
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
  if (h.lv) mg_free(&h);
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
