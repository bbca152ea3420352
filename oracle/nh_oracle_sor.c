/*
 * nh_oracle_sor.c -- CPU restatement of ORCTM's barotropic free-surface solve: the coefficients of
 * ITPREP (ORCTM/src_orctm/itprep.f90:15-38) and the right-hand side and red-black SOR iteration of
 * TRONEU (ORCTM/src_orctm/troneu.f90:48-131). TEST INFRASTRUCTURE ONLY: loaded by tests/ and
 * bench.py's CPU-baseline leg, never by the product. One process holds the whole domain
 * (p_ioff = p_joff = 0, have_g_* all true). Arrays are the caller's Fortran arrays with halos:
 *   p-points (IE,JE)          XX, FF, DLXP, DLYP, DTDXPO, DTDYO, WETO(:,:,1), B1O, Z1O
 *   u-points (I_start:IE,JE)  UF, DLXU, DLYU, DEUTO, U1O, UZO
 *   v-points (IE,J_start:JE)  VF, DLXV, DLYV, DEUTE, V1E, VZE
 * Every expression is evaluated left to right as the Fortran source writes it (no contraction).
 */
#include <stddef.h>

#include "nh_oracle_ocv.h"
#include "nh_oracle_sor.h"

#define P2(i, j) ((size_t)((i)-1) + (size_t)IE * (size_t)((j)-1))
#define U2(i, j) ((size_t)((i)-IS) + (size_t)NU * (size_t)((j)-1))
#define V2(i, j) ((size_t)((i)-1) + (size_t)IE * (size_t)((j)-JS))

/* itprep.f90:15-38: XX = DLXP*DLYP; UF = DEUTO*DLYU/DLXU*G*CONN*STABN*DT**2 (u-points);
 * VF = DEUTE*DLXV/DLYV*G*CONN*STABN*DT**2 (v-points); FF = 0, then for J = J_start+1..JE,
 * I = I_start+1..IE, FF = 1/(XX+UF(I)+UF(I-1)+VF(J)+VF(J-1)); bounds_exch(FF). */
void orc_itprep(const orc_sorgrid* g, double* xx, double* uf, double* vf, double* ff) {
  const int IE = g->IE, JE = g->JE, IS = g->i_start, JS = g->j_start;
  const int NU = IE - IS + 1;
  const double dt2 = g->dt * g->dt;
  for (int i = 1; i <= IE; ++i)
    for (int j = 1; j <= JE; ++j) xx[P2(i, j)] = g->dlxp[P2(i, j)] * g->dlyp[P2(i, j)];
  for (int i = IS; i <= IE; ++i)
    for (int j = 1; j <= JE; ++j)
      uf[U2(i, j)] = g->deuto[U2(i, j)] * g->dlyu[U2(i, j)] / g->dlxu[U2(i, j)] * g->g * g->conn * g->stabn * dt2;
  for (int i = 1; i <= IE; ++i)
    for (int j = JS; j <= JE; ++j)
      vf[V2(i, j)] = g->deute[V2(i, j)] * g->dlxv[V2(i, j)] / g->dlyv[V2(i, j)] * g->g * g->conn * g->stabn * dt2;
  for (int j = 1; j <= JE; ++j)
    for (int i = 1; i <= IE; ++i) ff[P2(i, j)] = 0.0;
  for (int j = JS + 1; j <= JE; ++j)
    for (int i = IS + 1; i <= IE; ++i)
      ff[P2(i, j)] = 1. / (xx[P2(i, j)] + uf[U2(i, j)] + uf[U2(i - 1, j)] + vf[V2(i, j)] + vf[V2(i, j - 1)]);
  orc_bounds_exch(ff, IE, JE, 1, 1, 1, g->cyc_x, g->cyc_y);
}

/* troneu.f90:48-131. b1o is read and written (only J = J_start+1..JE, I = I_start+1..IE are
 * recomputed, as in the source); z1o receives the result. */
void orc_troneu(const orc_sorgrid* g, const double* xx, const double* uf, const double* vf, const double* ff,
                const double* u1o, const double* uzo, const double* v1e, const double* vze, double* b1o,
                double* z1o) {
  const int IE = g->IE, JE = g->JE, IS = g->i_start, JS = g->j_start;
  const int NU = IE - IS + 1;
  const double cono = g->cono, conn = g->conn, sorpar = g->sorpar;
  for (int j = JS + 1; j <= JE; ++j)
    for (int i = IS + 1; i <= IE; ++i) {
      const double fx = g->dtdxpo[P2(i, j)] *
                        ((cono * u1o[U2(i - 1, j)] + uzo[U2(i - 1, j)] * conn) * g->dlyu[U2(i - 1, j)] -
                         (u1o[U2(i, j)] * cono + uzo[U2(i, j)] * conn) * g->dlyu[U2(i, j)]) /
                        g->dlyp[P2(i, j)];
      const double fy = g->dtdyo[P2(i, j)] *
                        ((v1e[V2(i, j)] * cono + conn * vze[V2(i, j)]) * g->dlxv[V2(i, j)] -
                         (v1e[V2(i, j - 1)] * cono + conn * vze[V2(i, j - 1)]) * g->dlxv[V2(i, j - 1)]) /
                        g->dlxp[P2(i, j)];
      const double b = g->weto1[P2(i, j)] * (fx + fy);
      b1o[P2(i, j)] = b * xx[P2(i, j)];
    }
  for (int j = 1; j <= JE; ++j)
    for (int i = 1; i <= IE; ++i) z1o[P2(i, j)] = 0.0;
  const double sorpai = 1. - sorpar;
  for (int j = 1; j <= JE; ++j) z1o[P2(1, j)] = b1o[P2(1, j)] / xx[P2(1, j)];
  for (int j = 1; j <= JE; ++j) z1o[P2(IE, j)] = b1o[P2(IE, j)] / xx[P2(IE, j)];
  for (int i = 1; i <= IE; ++i) z1o[P2(i, 1)] = b1o[P2(i, 1)] / xx[P2(i, 1)];
  for (int i = 1; i <= IE; ++i) z1o[P2(i, JE)] = b1o[P2(i, JE)] / xx[P2(i, JE)];
  for (int it = 0; it < g->iterations; ++it) {
    for (int color = 0; color < 2; ++color) {
      for (int j = 2; j <= JE - 1; ++j) {
        const int is = (j + color) % 2 + 2;
        for (int i = is; i <= IE - 1; i += 2) {
          const double s = b1o[P2(i, j)] + uf[U2(i, j)] * z1o[P2(i + 1, j)] + uf[U2(i - 1, j)] * z1o[P2(i - 1, j)] +
                           vf[V2(i, j)] * z1o[P2(i, j + 1)] + vf[V2(i, j - 1)] * z1o[P2(i, j - 1)];
          z1o[P2(i, j)] = sorpar * ff[P2(i, j)] * s + sorpai * z1o[P2(i, j)];
        }
      }
      orc_bounds_exch(z1o, IE, JE, 1, 1, 1, g->cyc_x, g->cyc_y);
    }
  }
}
