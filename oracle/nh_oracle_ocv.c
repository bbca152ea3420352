/*
 * nh_oracle_ocv.c -- CPU restatement of the steps of OCVTOT_NH on either side of the pressure solve
 * (ORCTM/src_orctm/ocvtot_nh.f90). TEST INFRASTRUCTURE ONLY: loaded by tests/ and never by the
 * product. Arrays are the caller's Fortran arrays with halos (mo_commo1.f90:105-146):
 *   p-points (IE,JE,KE)            PNH, DDPO, WETO, DIVG          F(i,j,k) = a[(i-1) + IE*((j-1) + JE*(k-1))]
 *   u-points (I_start:IE,JE,KE)    UOO, UOO_1, AMSUO, DDUO
 *   v-points (IE,J_start:JE,KE)    VOE, VOE_1, AMSUE, DDUE
 *   w-points (IE,JE,KEP)           WNO, WNO_1
 *   2-D (IE,JE) DLXP, DLYP, DTDXPO, DTDYO, ZO; (I_start:IE,JE) DTDXUO, DLYU; (IE,J_start:JE) DPYE, DLXV.
 * One process holds the whole domain (nprocx = nprocy = 1, have_g_* all true). Every expression is
 * evaluated left to right as written in the Fortran source.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#include "nh_oracle_ocv.h"

#define P3(i, j, k) ((size_t)((i)-1) + (size_t)IE * ((size_t)((j)-1) + (size_t)JE * (size_t)((k)-1)))
#define U3(i, j, k) ((size_t)((i)-IS) + (size_t)NU * ((size_t)((j)-1) + (size_t)JE * (size_t)((k)-1)))
#define V3(i, j, k) ((size_t)((i)-1) + (size_t)IE * ((size_t)((j)-JS) + (size_t)NV * (size_t)((k)-1)))
#define P2(i, j) ((size_t)((i)-1) + (size_t)IE * (size_t)((j)-1))
#define U2(i, j) ((size_t)((i)-IS) + (size_t)NU * (size_t)((j)-1))
#define V2(i, j) ((size_t)((i)-1) + (size_t)IE * (size_t)((j)-JS))

/*
 * bounds_exch_3d (mo_parallel.f90:551-668) for one process. x: with ICYCLI_X, a(1) = a(IE-1) and
 * a(IE) = a(2) (the values taken before either copy), for every j and k of the array. y, after x:
 * with ICYCLI_Y, for arrays whose j range starts at 0 (v-points with J_start = 0) a(0) = a(JE-1) and
 * a(JE) = a(1), otherwise a(1) = a(JE-1) and a(JE) = a(2). Without the cyclic flags nothing changes.
 * a(i0..IE, j0..JE, 1..nk) with i0, j0 in {0, 1}.
 */
void orc_bounds_exch(double* a, int IE, int JE, int nk, int i0, int j0, int cyc_x, int cyc_y) {
  const int ni = IE - i0 + 1, nj = JE - j0 + 1;
#define A(i, j, k) a[(size_t)((i)-i0) + (size_t)ni * ((size_t)((j)-j0) + (size_t)nj * (size_t)((k)-1))]
  if (cyc_x) {
    for (int k = 1; k <= nk; ++k)
      for (int j = j0; j <= JE; ++j) {
        const double xs1 = A(2, j, k), xs2 = A(IE - 1, j, k);
        A(1, j, k) = xs2;
        A(IE, j, k) = xs1;
      }
  }
  if (cyc_y) {
    for (int k = 1; k <= nk; ++k)
      for (int i = i0; i <= IE; ++i) {
        const double y1 = (j0 == 0) ? A(i, 1, k) : A(i, 2, k), y2 = A(i, JE - 1, k);
        if (j0 == 0)
          A(i, 0, k) = y2;
        else
          A(i, 1, k) = y2;
        A(i, JE, k) = y1;
      }
  }
#undef A
}

/*
 * ocvtot_nh.f90:11-62 -- GU = (UOO-UOO_1)/DT, GV = (VOE-VOE_1)/DT, GW = (WNO-WNO_1)/DT;
 * DIVGU = (GU(I)-GU(I-1))/DLXP, DIVGV = (GV(J-1)-GV(J))/DLYP,
 * DIVGW = (GW(K)-GW(K+1))/max(DDPO,1e-3) (K = 1: /(max(DDPO,1e-3)+ZO)), DIVG = DIVGU+DIVGV+DIVGW.
 * Written for the interior cells i = 2..IE-1, j = 2..JE-1, all k (the only cells ComputeRHS reads,
 * mo_petsc.f90:376); the halo of divg is left as it is.
 */
void orc_divergence(const orc_vgrid* g, const double* uoo, const double* uoo_1, const double* voe,
                    const double* voe_1, const double* wno, const double* wno_1, const double* zo, double* divg) {
  const int IE = g->IE, JE = g->JE, KE = g->KE, IS = g->i_start, JS = g->j_start;
  const int NU = IE - IS + 1, NV = JE - JS + 1;
  const double dt = g->dt;
  for (int k = 1; k <= KE; ++k)
    for (int j = 2; j <= JE - 1; ++j)
      for (int i = 2; i <= IE - 1; ++i) {
        const double gu_i = (uoo[U3(i, j, k)] - uoo_1[U3(i, j, k)]) / dt;
        const double gu_m = (uoo[U3(i - 1, j, k)] - uoo_1[U3(i - 1, j, k)]) / dt;
        const double gv_j = (voe[V3(i, j, k)] - voe_1[V3(i, j, k)]) / dt;
        const double gv_m = (voe[V3(i, j - 1, k)] - voe_1[V3(i, j - 1, k)]) / dt;
        const double gw_k = (wno[P3(i, j, k)] - wno_1[P3(i, j, k)]) / dt;
        const double gw_p = (wno[P3(i, j, k + 1)] - wno_1[P3(i, j, k + 1)]) / dt;
        const double divgu = (gu_i - gu_m) / g->dlxp[P2(i, j)];
        const double divgv = (gv_m - gv_j) / g->dlyp[P2(i, j)];
        const double h = fmax(g->ddpo[P3(i, j, k)], 1.0e-3);
        const double divgw = (k == 1) ? (gw_k - gw_p) / (h + zo[P2(i, j)]) : (gw_k - gw_p) / h;
        divg[P3(i, j, k)] = divgu + divgv + divgw;
      }
}

/*
 * ocvtot_nh.f90:151-236 after the solve, for one process:
 *   bounds_exch(PNH) (:151); open-boundary copies of PNH (:154-186);
 *   UOO += AMSUO*DTDXUO*(PNH(I)-PNH(I+1)), J = 2..JE1, I = I_start+1..IE1 (:189-194); bounds_exch(UOO);
 *   VOE += AMSUE*DPYE*(PNH(J+1)-PNH(J)), J = J_start+1..JE1, I = 2..IE1 (:196-201); bounds_exch(VOE);
 *   WNO = 0 (:207-211); WNO(K) = WNO(K+1) + DTI*WETO*(DTDXPO*(UOO(I-1)*DDUO(I-1)*DLYU(I-1)
 *     - UOO(I)*DDUO(I)*DLYU(I))/DLYP) + DTI*WETO*(+DTDYO*(VOE(J)*DDUE(J)*DLXV(J)
 *     - VOE(J-1)*DDUE(J-1)*DLXV(J-1))/DLXP), K = KE..1, J = J_start+1..JE, I = I_start+1..IE (:213-228);
 *   bounds_exch(WNO) (:230).
 * pnh holds the solution in its interior on entry (its halo as the caller left it); pnh, uoo, voe
 * and wno are updated in place. The copies UOO_1 = UOO, ... (:232-249) stay with the caller.
 */
void orc_project(const orc_vgrid* g, double* pnh, double* uoo, double* voe, double* wno) {
  const int IE = g->IE, JE = g->JE, KE = g->KE, IS = g->i_start, JS = g->j_start;
  const int NU = IE - IS + 1, NV = JE - JS + 1, KEP = KE + 1;
  orc_bounds_exch(pnh, IE, JE, KE, 1, 1, g->cyc_x, g->cyc_y);
  if (!g->cyc_x) {
    if (g->open_w)
      for (int j = 2; j <= JE - 1; ++j)
        for (int k = 1; k <= KE; ++k) pnh[P3(1, j, k)] = pnh[P3(2, j, k)];
    if (g->open_e)
      for (int j = 2; j <= JE - 1; ++j)
        for (int k = 1; k <= KE; ++k) pnh[P3(IE, j, k)] = pnh[P3(IE - 1, j, k)];
  }
  if (!g->cyc_y) {
    if (g->open_n)
      for (int i = 1; i <= IE; ++i)
        for (int k = 1; k <= KE; ++k) pnh[P3(i, 1, k)] = pnh[P3(i, 2, k)];
    if (g->open_s)
      for (int i = 1; i <= IE; ++i)
        for (int k = 1; k <= KE; ++k) pnh[P3(i, JE, k)] = pnh[P3(i, JE - 1, k)];
  }
  for (int k = 1; k <= KE; ++k)
    for (int j = 2; j <= JE - 1; ++j)
      for (int i = IS + 1; i <= IE - 1; ++i)
        uoo[U3(i, j, k)] = uoo[U3(i, j, k)] +
                           g->amsuo[U3(i, j, k)] * g->dtdxuo[U2(i, j)] * (pnh[P3(i, j, k)] - pnh[P3(i + 1, j, k)]);
  orc_bounds_exch(uoo, IE, JE, KE, IS, 1, g->cyc_x, g->cyc_y);
  for (int k = 1; k <= KE; ++k)
    for (int j = JS + 1; j <= JE - 1; ++j)
      for (int i = 2; i <= IE - 1; ++i)
        voe[V3(i, j, k)] = voe[V3(i, j, k)] +
                           g->amsue[V3(i, j, k)] * g->dpye[V2(i, j)] * (pnh[P3(i, j + 1, k)] - pnh[P3(i, j, k)]);
  orc_bounds_exch(voe, IE, JE, KE, 1, JS, g->cyc_x, g->cyc_y);
  memset(wno, 0, sizeof(double) * (size_t)IE * JE * KEP);
  for (int k = KE; k >= 1; --k)
    for (int j = JS + 1; j <= JE; ++j)
      for (int i = IS + 1; i <= IE; ++i) {
        const double dw = g->dti * g->weto[P3(i, j, k)];
        const double fx = g->dtdxpo[P2(i, j)] *
                          (uoo[U3(i - 1, j, k)] * g->dduo[U3(i - 1, j, k)] * g->dlyu[U2(i - 1, j)] -
                           uoo[U3(i, j, k)] * g->dduo[U3(i, j, k)] * g->dlyu[U2(i, j)]) /
                          g->dlyp[P2(i, j)];
        const double fy = g->dtdyo[P2(i, j)] *
                          (voe[V3(i, j, k)] * g->ddue[V3(i, j, k)] * g->dlxv[V2(i, j)] -
                           voe[V3(i, j - 1, k)] * g->ddue[V3(i, j - 1, k)] * g->dlxv[V2(i, j - 1)]) /
                          g->dlxp[P2(i, j)];
        wno[P3(i, j, k)] = wno[P3(i, j, k + 1)] + dw * fx + dw * fy;
      }
  orc_bounds_exch(wno, IE, JE, KEP, 1, 1, g->cyc_x, g->cyc_y);
}
