/*
 * nh_oracle.c -- CPU ORACLE for the nonhydrostatic pressure Poisson path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle_nh.so, and only to
 * check (or, for cpu_baseline, to time beside) the HIP library. The product path
 * (orctm_amd/, liborctm_nh.so) never links, loads or calls anything here.
 *
 * What it restates (all FP64 unless noted, file:line relative to the reference tree):
 *   orc_assemble      compute_matrix, ORCTM/src_orctm/mo_petsc.f90:543-782 (operator values of the
 *                     MatSetValuesStencil calls :626,:765), incl. the ZW/ZP pre-pass :557-590
 *   orc_rhs           ComputeRHS, mo_petsc.f90:298-376 (active statement :376, b = DIVG*WETO)
 *   orc_matvec64      the AIJ MatMult KSPSolve applies to that matrix (PETSc internal); here as a
 *                     fixed-order fma chain centre,W,E,N,S,U,D which the device kernel also uses
 *   orc_compact       the compact form of the same operator (per-cell ZW/ZP + flag bits + per-column
 *                     metric products) the device keeps; orc_rebuild64 rebuilds compute_matrix's values
 *                     from it bit for bit, orc_inner_coef32_compact evaluates the FP32 inner operator
 *   orc_inner_*       the scaled operator C = S*(-A), S = DLXP*DLYP/ZP, in FP64 or FP32, its
 *                     vertical-line preconditioner and the PCG/refinement driver. These have no
 *                     counterpart in the reference (which runs FGMRES + PCMG inside PETSc,
 *                     ORCTM/verification/Sect.3.1/restart_nonhydro.sh:30); they restate the
 *                     algorithm the HIP library runs so that it can be checked on the CPU.
 *   orc_fgmres        FGMRES(m), right-preconditioned, FP64: the reference's Krylov method
 *                     (-cg3d_ksp_type fgmres, restart 30 = PETSc default) with the line
 *                     preconditioner standing in for PCMG(1 level) [PETSc internals unverified].
 *
 * How each part is pinned (tests/test_oracle_*.py): closed forms on uniform grids (separable
 * operator, DCT eigenvectors in x/y, hand-derived vertical stencil), worked examples with literal
 * values for partial cells and the one-cell-deep column, row-sum / land-row / symmetry invariants,
 * an independent dense transliteration on tiny grids, SciPy direct solves (golden fixtures) and
 * scipy.linalg.solve_banded for the line solve. The reference ships no golden vectors of its own.
 *
 * Interior vectors use DMDA natural order: n = i + nx*(j + ny*k), i fastest (mo_petsc.f90:620-622).
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/orctm_nh.h"

#define NC 7 /* coefficient planes: 0 centre, 1 west(i-1), 2 east(i+1), 3 north(j-1), 4 south(j+1),
                5 up(k-1), 6 down(k+1) */

/* Wall clock in milliseconds (the solvers report the time of their iterations apart from their set-up). */
static double wall_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

/* Fortran element accessors (1-based indices, column-major). */
#define P2(a, I, J) ((a)[((size_t)(I)-1) + (size_t)IE * ((size_t)(J)-1)])
#define U2(a, I, J) ((a)[((size_t)(I) - (size_t)g->i_start) + (size_t)(IE - g->i_start + 1) * ((size_t)(J)-1)])
#define V2(a, I, J) ((a)[((size_t)(I)-1) + (size_t)IE * ((size_t)(J) - (size_t)g->j_start)])
#define P3(a, I, J, K) ((a)[((size_t)(I)-1) + (size_t)IE * (((size_t)(J)-1) + (size_t)JE * ((size_t)(K)-1))])

int orc_abi_version(void) { return 3; }

static int grid_ok(const orctm_nh_grid_t* g) {
  if (!g) return 0;
  if (g->ie < 3 || g->je < 3 || g->ke < 1) return 0;
  if (g->i_start < 0 || g->i_start > 1 || g->j_start < 0 || g->j_start > 1) return 0;
  if (!g->dlxp || !g->dlyp || !g->dlxu || !g->dlyv || !g->ddpo || !g->weto || !g->dzw) return 0;
  return 1;
}

/* ZW(K), K=1..KEP and ZP(K), K=1..KE of one column (I,J): mo_petsc.f90:557-590. */
static void column_zw_zp(const orctm_nh_grid_t* g, int I, int J, double* zw, double* zp) {
  const int IE = g->ie, JE = g->je, KE = g->ke, KEP = KE + 1;
  for (int K = 1; K <= KEP; ++K) {
    double dd;
    if (K == 1) {
      dd = (P3(g->ddpo, I, J, 1) == 0.0) ? g->dzw[0] / 2.0 : P3(g->ddpo, I, J, 1) / 2.0;
    } else if (K == KEP) {
      dd = (P3(g->ddpo, I, J, KE) == 0.0) ? g->dzw[KE - 1] : P3(g->ddpo, I, J, KE);
    } else {
      const double dm = P3(g->ddpo, I, J, K - 1), d0 = P3(g->ddpo, I, J, K);
      if (dm == 0.0)
        dd = (g->dzw[K - 1] + g->dzw[K - 2]) / 2.0;
      else if (d0 == 0.0)
        dd = dm;
      else
        dd = (dm + d0) / 2.0;
    }
    zw[K - 1] = 1.0 / dd;
  }
  for (int K = 1; K <= KE; ++K) zp[K - 1] = 1.0 / ((1.0 / zw[K - 1] + 1.0 / zw[K]) / 2.0);
}

/*
 * compute_matrix (mo_petsc.f90:543-782). A: NC planes of N doubles (see NC). S (may be NULL): the
 * per-cell weight DLXP*DLYP/ZP that makes the vertical part of -A symmetric.
 * Returns ORCTM_NH_OK or ORCTM_NH_ERR_INVALID (bad grid, non-positive spacing or non-finite ZW/ZP).
 */
int orc_assemble(const orctm_nh_grid_t* g, double* A, double* S) {
  if (!grid_ok(g) || !A) return ORCTM_NH_ERR_INVALID;
  const int IE = g->ie, JE = g->je, KE = g->ke;
  const int IE1 = IE - 1, JE1 = JE - 1;
  const int nx = IE - 2, ny = JE - 2, nz = KE;
  const size_t N = (size_t)nx * ny * nz;
  int bad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(| : bad)
  for (int J = 2; J <= JE1; ++J) {
    for (int I = 2; I <= IE1; ++I) {
      double* zw = (double*)malloc(sizeof(double) * (2 * (size_t)KE + 1));
      double* zp = zw + KE + 1;
      column_zw_zp(g, I, J, zw, zp);
      const double dlxuw = U2(g->dlxu, I - 1, J), dlxue = U2(g->dlxu, I, J);
      const double dlyvn = V2(g->dlyv, I, J - 1), dlyvs = V2(g->dlyv, I, J);
      const double dlxp = P2(g->dlxp, I, J), dlyp = P2(g->dlyp, I, J);
      if (!(dlxuw > 0) || !(dlxue > 0) || !(dlyvn > 0) || !(dlyvs > 0) || !(dlxp > 0) || !(dlyp > 0)) bad = 1;
      const double xuw = 1.0 / dlxuw, xue = 1.0 / dlxue, xp = 1.0 / dlxp;
      const double yvn = 1.0 / dlyvn, yvs = 1.0 / dlyvs, yp = 1.0 / dlyp;
      for (int K = 1; K <= KE; ++K) {
        if (!isfinite(zw[K - 1]) || !isfinite(zw[K]) || !isfinite(zp[K - 1]) || !(zp[K - 1] > 0)) bad = 1;
        double X_W = xuw * xp, X_C = (xuw + xue) * xp, X_E = xue * xp;
        double Y_N = yvn * yp, Y_C = (yvn + yvs) * yp, Y_S = yvs * yp;
        double Z_U = zw[K - 1] * zp[K - 1], Z_C = (zw[K - 1] + zw[K]) * zp[K - 1], Z_D = zw[K] * zp[K - 1];
        double c[NC] = {0, 0, 0, 0, 0, 0, 0};
        if (P3(g->weto, I, J, K) == 0.0) {
          /* land row: diagonal only, unmodified terms (mo_petsc.f90:623-626) */
          c[0] = -(X_C + Y_C + Z_C);
        } else {
          const double ww = P3(g->weto, I - 1, J, K), we = P3(g->weto, I + 1, J, K);
          const double wn = P3(g->weto, I, J - 1, K), ws = P3(g->weto, I, J + 1, K);
          /* homogeneous Neumann at land, mo_petsc.f90:630-653 */
          if (ww == 0.0 && we == 1.0) {
            X_W = 0.0;
            X_C = xue * xp;
          } else if (ww == 1.0 && we == 0.0) {
            X_E = 0.0;
            X_C = xuw * xp;
          } else if (ww == 0.0 && we == 0.0) {
            X_W = 0.0;
            X_E = 0.0;
            X_C = 0.0;
          }
          if (wn == 0.0 && ws == 1.0) {
            Y_N = 0.0;
            Y_C = yvs * yp;
          } else if (wn == 1.0 && ws == 0.0) {
            Y_S = 0.0;
            Y_C = yvn * yp;
          } else if (wn == 0.0 && ws == 0.0) {
            Y_N = 0.0;
            Y_S = 0.0;
            Y_C = 0.0;
          }
          /* surface Dirichlet / bottom Neumann, mo_petsc.f90:657-665 (IF/ELSEIF order kept: at K=1
             the bottom test is never made) */
          if (K == 1) {
            Z_U = 0.0;
          } else if (K == KE) {
            Z_D = 0.0;
            Z_C = zw[K - 1] * zp[K - 1];
          } else if (P3(g->weto, I, J, K + 1) == 0.0) {
            Z_D = 0.0;
            Z_C = zw[K - 1] * zp[K - 1];
          }
          /* open boundaries, mo_petsc.f90:669-689 (centre keeps the term: Dirichlet 0) */
          if (!g->cyclic_x) {
            if (g->open_w && I == 2) X_W = 0.0;
            if (g->open_e && I == IE1) X_E = 0.0;
          }
          if (!g->cyclic_y) {
            if (g->open_n && J == 2) Y_N = 0.0;
            if (g->open_s && J == JE1) Y_S = 0.0;
          }
          c[0] = -(X_C + Y_C + Z_C);
          /* couplings inserted only to wet cells inside the DMDA (mo_petsc.f90:692-764) */
          if (ww == 1.0 && (g->cyclic_x || I != 2)) c[1] = X_W;
          if (we == 1.0 && (g->cyclic_x || I != IE1)) c[2] = X_E;
          if (wn == 1.0 && (g->cyclic_y || J != 2)) c[3] = Y_N;
          if (ws == 1.0 && (g->cyclic_y || J != JE1)) c[4] = Y_S;
          if (K != 1) c[5] = Z_U;
          if (K != KE && P3(g->weto, I, J, K + 1) == 1.0) c[6] = Z_D;
        }
        const size_t n = (size_t)(I - 2) + (size_t)nx * ((size_t)(J - 2) + (size_t)ny * (size_t)(K - 1));
        for (int q = 0; q < NC; ++q) A[(size_t)q * N + n] = c[q];
        if (S) S[n] = (dlxp * dlyp) / zp[K - 1];
      }
      free(zw);
    }
  }
  return bad ? ORCTM_NH_ERR_INVALID : ORCTM_NH_OK;
}

/* ------------------------------------------------------------------------------------------- */
/* Compact form of the same operator: everything compute_matrix writes is fixed by per-cell ZW/ZP,
   12 flag bits and 7 per-column metric products. The device keeps only this form and rebuilds each
   row on the fly; orc_row64 rebuilds it here with the operations orc_assemble uses.
     flags bit 0     WETO(I,J,K) == 1
           bits 1-4  coupling present W, E, N, S   (mo_petsc.f90:692-764)
           bit 5     coupling present U (K != 1)
           bit 6     coupling present D (K != KE and WETO(K+1) == 1)
           bits 7-8  X_C: 0 (XUw+XUe)*XP, 1 XUe*XP, 2 XUw*XP, 3 zero  (mo_petsc.f90:630-640)
           bits 9-10 Y_C: 0 (YVn+YVs)*YP, 1 YVs*YP, 2 YVn*YP, 3 zero  (mo_petsc.f90:643-653)
           bit 11    Z_C: 0 (ZW(K)+ZW(K+1))*ZP(K), 1 ZW(K)*ZP(K)      (mo_petsc.f90:657-665)
     col planes 0-6  XUw*XP, XUe*XP, YVn*YP, YVs*YP, (XUw+XUe)*XP, (YVn+YVs)*YP, DLXP*DLYP        */

int orc_compact(const orctm_nh_grid_t* g, double* zw, double* zp, uint16_t* flags, double* col) {
  if (!grid_ok(g) || !zw || !zp || !flags || !col) return ORCTM_NH_ERR_INVALID;
  const int IE = g->ie, JE = g->je, KE = g->ke;
  const int IE1 = IE - 1, JE1 = JE - 1;
  const int nx = IE - 2, ny = JE - 2;
  const size_t P = (size_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int J = 2; J <= JE1; ++J) {
    for (int I = 2; I <= IE1; ++I) {
      const size_t c = (size_t)(I - 2) + (size_t)nx * (J - 2);
      double* zwc = (double*)malloc(sizeof(double) * (2 * (size_t)KE + 1));
      double* zpc = zwc + KE + 1;
      column_zw_zp(g, I, J, zwc, zpc);
      const double xuw = 1.0 / U2(g->dlxu, I - 1, J), xue = 1.0 / U2(g->dlxu, I, J), xp = 1.0 / P2(g->dlxp, I, J);
      const double yvn = 1.0 / V2(g->dlyv, I, J - 1), yvs = 1.0 / V2(g->dlyv, I, J), yp = 1.0 / P2(g->dlyp, I, J);
      col[0 * P + c] = xuw * xp;
      col[1 * P + c] = xue * xp;
      col[2 * P + c] = yvn * yp;
      col[3 * P + c] = yvs * yp;
      col[4 * P + c] = (xuw + xue) * xp;
      col[5 * P + c] = (yvn + yvs) * yp;
      col[6 * P + c] = P2(g->dlxp, I, J) * P2(g->dlyp, I, J);
      for (int K = 1; K <= KE + 1; ++K) zw[(size_t)(K - 1) * P + c] = zwc[K - 1];
      for (int K = 1; K <= KE; ++K) {
        zp[(size_t)(K - 1) * P + c] = zpc[K - 1];
        unsigned f = 0;
        if (P3(g->weto, I, J, K) != 0.0) {
          const double ww = P3(g->weto, I - 1, J, K), we = P3(g->weto, I + 1, J, K);
          const double wn = P3(g->weto, I, J - 1, K), ws = P3(g->weto, I, J + 1, K);
          f |= 1u;
          if (ww == 1.0 && (g->cyclic_x || I != 2)) f |= 1u << 1;
          if (we == 1.0 && (g->cyclic_x || I != IE1)) f |= 1u << 2;
          if (wn == 1.0 && (g->cyclic_y || J != 2)) f |= 1u << 3;
          if (ws == 1.0 && (g->cyclic_y || J != JE1)) f |= 1u << 4;
          if (K != 1) f |= 1u << 5;
          if (K != KE && P3(g->weto, I, J, K + 1) == 1.0) f |= 1u << 6;
          unsigned xc = 0, yc = 0, zc = 0;
          if (ww == 0.0 && we == 1.0) xc = 1;
          else if (ww == 1.0 && we == 0.0) xc = 2;
          else if (ww == 0.0 && we == 0.0) xc = 3;
          if (wn == 0.0 && ws == 1.0) yc = 1;
          else if (wn == 1.0 && ws == 0.0) yc = 2;
          else if (wn == 0.0 && ws == 0.0) yc = 3;
          if (K == 1) zc = 0;
          else if (K == KE || P3(g->weto, I, J, K + 1) == 0.0) zc = 1;
          f |= xc << 7 | yc << 9 | zc << 11;
        }
        flags[(size_t)(K - 1) * P + c] = (uint16_t)f;
      }
      free(zwc);
    }
  }
  return ORCTM_NH_OK;
}

/* One row of the operator from the compact form, in REAL arithmetic (double: bit-exact with
   orc_assemble; float: the inner operator's arithmetic). c7 receives centre, W, E, N, S, U, D. */
#define ORC_ROW(REAL, NAME)                                                                                 \
  static void NAME(const REAL* colc, REAL zwk, REAL zwk1, REAL zpk, unsigned f, REAL* c7) {                 \
    const REAL xw = colc[0], xe = colc[1], yn = colc[2], ys = colc[3];                                      \
    const unsigned xc = (f >> 7) & 3u, yc = (f >> 9) & 3u;                                                  \
    const REAL X_C = xc == 0 ? colc[4] : xc == 1 ? xe : xc == 2 ? xw : (REAL)0;                             \
    const REAL Y_C = yc == 0 ? colc[5] : yc == 1 ? ys : yc == 2 ? yn : (REAL)0;                             \
    const REAL Z_C = ((f >> 11) & 1u) ? zwk * zpk : (zwk + zwk1) * zpk;                                     \
    c7[0] = -(X_C + Y_C + Z_C);                                                                             \
    c7[1] = (f & (1u << 1)) ? xw : (REAL)0;                                                                 \
    c7[2] = (f & (1u << 2)) ? xe : (REAL)0;                                                                 \
    c7[3] = (f & (1u << 3)) ? yn : (REAL)0;                                                                 \
    c7[4] = (f & (1u << 4)) ? ys : (REAL)0;                                                                 \
    c7[5] = (f & (1u << 5)) ? zwk * zpk : (REAL)0;                                                          \
    c7[6] = (f & (1u << 6)) ? zwk1 * zpk : (REAL)0;                                                         \
  }
ORC_ROW(double, row64)
ORC_ROW(float, row32)

/* The reference operator rebuilt from the compact form (must equal orc_assemble bit for bit). */
void orc_rebuild64(int nx, int ny, int nz, const double* zw, const double* zp, const uint16_t* flags,
                   const double* col, double* A) {
  const size_t P = (size_t)nx * ny, N = P * nz;
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < P; ++c) {
    double cc[7], c7[7];
    for (int q = 0; q < 7; ++q) cc[q] = col[(size_t)q * P + c];
    for (int k = 0; k < nz; ++k) {
      const size_t n = c + P * k;
      row64(cc, zw[n], zw[n + P], zp[n], flags[n], c7);
      for (int q = 0; q < 7; ++q) A[(size_t)q * N + n] = c7[q];
    }
  }
}

/* The FP32 inner operator C = S*(-A), S = DLXP*DLYP/ZP, evaluated in float from the compact form
   rounded to float (what the device matvec computes on the fly). */
void orc_inner_coef32_compact(int nx, int ny, int nz, const double* zw, const double* zp, const uint16_t* flags,
                              const double* col, float* C) {
  const size_t P = (size_t)nx * ny, N = P * nz;
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < P; ++c) {
    float cc[7], c7[7];
    for (int q = 0; q < 7; ++q) cc[q] = (float)col[(size_t)q * P + c];
    for (int k = 0; k < nz; ++k) {
      const size_t n = c + P * k;
      const float zpk = (float)zp[n];
      row32(cc, (float)zw[n], (float)zw[n + P], zpk, flags[n], c7);
      const float S = cc[6] / zpk;
      for (int q = 0; q < 7; ++q) C[(size_t)q * N + n] = S * (-c7[q]);
    }
  }
}

/* ComputeRHS, mo_petsc.f90:298-376: b = DIVG*WETO on the interior. */
int orc_rhs(const orctm_nh_grid_t* g, const double* divg, double* b) {
  if (!grid_ok(g) || !divg || !b) return ORCTM_NH_ERR_INVALID;
  const int IE = g->ie, JE = g->je, KE = g->ke;
  const int nx = IE - 2, ny = JE - 2;
  for (int K = 1; K <= KE; ++K)
    for (int J = 2; J <= JE - 1; ++J)
      for (int I = 2; I <= IE - 1; ++I)
        b[(size_t)(I - 2) + (size_t)nx * ((size_t)(J - 2) + (size_t)ny * (K - 1))] =
            P3(divg, I, J, K) * P3(g->weto, I, J, K);
  return ORCTM_NH_OK;
}

/* Copy-back of mo_petsc.f90:223-227: interior vector -> PNH(2:IE1,2:JE1,1:KE), halo untouched. */
int orc_scatter_interior(const orctm_nh_grid_t* g, const double* x, double* pnh) {
  if (!grid_ok(g) || !x || !pnh) return ORCTM_NH_ERR_INVALID;
  const int IE = g->ie, JE = g->je, KE = g->ke;
  const int nx = IE - 2, ny = JE - 2;
  for (int K = 1; K <= KE; ++K)
    for (int J = 2; J <= JE - 1; ++J)
      for (int I = 2; I <= IE - 1; ++I)
        P3(pnh, I, J, K) = x[(size_t)(I - 2) + (size_t)nx * ((size_t)(J - 2) + (size_t)ny * (K - 1))];
  return ORCTM_NH_OK;
}

/* ------------------------------------------------------------------------------------------- */
/* Precision-generic pieces: matvec, line preconditioner, PCG. Instantiated for double and float. */

/*
 * Storage of the FP32 line-preconditioner factors (device: nh_kernels.hip, k_line_factor<float>).
 * Each value is kept as an IEEE fp16 (10-bit mantissa, round to nearest even), only in the normal
 * fp16 range: l with |l| < 2^-14 is flushed to 0; invD is scaled per column by 2^sc, sc = 15 - ex
 * where max_k |invD| = f 2^ex, f in [0.5, 1), so the column's largest scaled value lies in
 * [2^14, 2^15); scaled magnitudes below 2^-14 are raised to 2^-14 (invD stays > 0, M stays SPD).
 * half_round returns the fp16 value as a float (exactly representable); the scale 2^-sc is exact.
 */
static float half_round(float v) {
  if (!(fabsf(v) < 65520.0f)) return isnan(v) ? v : copysignf(INFINITY, v);
  if (v == 0.0f) return v;
  uint32_t u;
  memcpy(&u, &v, 4);
  u = (u + 0xFFFu + ((u >> 13) & 1u)) & 0xFFFFE000u; /* 23 -> 10 mantissa bits, ties to even */
  memcpy(&v, &u, 4);
  return v;
}

static void factor_round64(double* invD, double* L, size_t P, int nz, size_t col) {
  (void)invD, (void)L, (void)P, (void)nz, (void)col; /* FP64 factors are kept as they are */
}

static void factor_round32(float* invD, float* L, size_t P, int nz, size_t col) {
  const float lo = 6.103515625e-05f; /* 2^-14, the smallest normal fp16 */
  float mx = 0.0f;
  for (int k = 0; k < nz; ++k) {
    const float a = fabsf(invD[col + P * (size_t)k]);
    if (a > mx) mx = a;
  }
  int ex = 0;
  if (mx > 0.0f && isfinite(mx)) frexpf(mx, &ex);
  const int sc = (mx > 0.0f && isfinite(mx)) ? 15 - ex : 0;
  for (int k = 0; k < nz; ++k) {
    const size_t n = col + P * (size_t)k;
    float v = ldexpf(invD[n], sc);
    if (fabsf(v) < lo) v = copysignf(lo, v);
    invD[n] = ldexpf(half_round(v), -sc);
    const float l = L[n];
    L[n] = fabsf(l) < lo ? 0.0f : half_round(l);
  }
}

#define REAL double
#define SFX(name) name##64
#define FMA fma
#include "nh_oracle_krylov.inc"
#undef REAL
#undef SFX
#undef FMA

#define REAL float
#define SFX(name) name##32
#define FMA fmaf
#include "nh_oracle_krylov.inc"
#undef REAL
#undef SFX
#undef FMA

/* ------------------------------------------------------------------------------------------- */
/* FP64 iterative refinement around the inner PCG (the algorithm of the HIP library).           */

static double dot64(size_t n, const double* a, const double* b) {
  /* deterministic: fixed chunks summed in order */
  const size_t CH = 4096;
  const size_t nch = (n + CH - 1) / CH;
  double* part = (double*)malloc(sizeof(double) * (nch ? nch : 1));
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < nch; ++c) {
    double s = 0.0;
    const size_t e = (c + 1) * CH < n ? (c + 1) * CH : n;
    for (size_t i = c * CH; i < e; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (size_t c = 0; c < nch; ++c) s += part[c];
  free(part);
  return s;
}

/*
 * Full solve: A p = b with FP64 refinement. precision: ORCTM_NH_FP64 or ORCTM_NH_MIXED.
 * A, S from orc_assemble. p: in = initial guess, out = solution (interior, natural order).
 * Returns a status; st (may be NULL) receives iterations and the final relative residual.
 */
int orc_solve(int nx, int ny, int nz, int cyc_x, int cyc_y, const double* A, const double* S, const double* b,
              double* p, int precision, double rtol, double rtol_inner, int maxit, int max_outer,
              orctm_nh_stats_t* st, const void* Cin) {
  const double t_begin = wall_ms();
  const size_t N = (size_t)nx * ny * nz;
  orctm_nh_stats_t s;
  memset(&s, 0, sizeof s);
  int status = ORCTM_NH_OK;
  const double nb = sqrt(dot64(N, b, b));
  s.bnorm = nb;
  if (!isfinite(nb)) {
    s.status = ORCTM_NH_ERR_NONFINITE;
    if (st) *st = s;
    return ORCTM_NH_ERR_NONFINITE;
  }
  if (nb == 0.0) { /* b = 0 -> p = 0 (A is nonsingular: Dirichlet surface) */
    memset(p, 0, sizeof(double) * N);
    if (st) *st = s;
    return ORCTM_NH_OK;
  }
  double* r = (double*)malloc(sizeof(double) * N);
  void* C = NULL;
  void* f1 = NULL;
  void* f2 = NULL;
  void* rin = NULL;
  void* x = NULL;
  void* work = NULL;
  const size_t es = precision == ORCTM_NH_FP64 ? sizeof(double) : sizeof(float);
  C = malloc(es * NC * N);
  f1 = malloc(es * N);
  f2 = malloc(es * N);
  rin = malloc(es * N);
  x = malloc(es * N);
  work = malloc(es * 4 * N);
  if (Cin) memcpy(C, Cin, es * NC * N);
  if (precision == ORCTM_NH_FP64) {
    if (!Cin) orc_inner_coef64(N, A, S, (double*)C);
    orc_line_factor64(nx, ny, nz, (const double*)C, (double*)f1, (double*)f2);
  } else {
    if (!Cin) orc_inner_coef32(N, A, S, (float*)C);
    orc_line_factor32(nx, ny, nz, (const float*)C, (float*)f1, (float*)f2);
  }
  /* the inner right-hand side is normalised: rin = S.*(-r) * gamma, gamma = 1/(max S * ||r_prev||),
     so that FP32 sums of squares stay far from overflow; the update undoes the scale */
  double smax = 0.0;
  for (size_t i = 0; i < N; ++i)
    if (S[i] > smax) smax = S[i];
  double nr_prev = nb;
  int inner_total = 0;
  int m = 0;
  double rel = 0.0;
  const double t_iter = wall_ms(); /* set-up (inner operator, line factors, work arrays) done */
  for (;; ++m) {
    orc_matvec64(nx, ny, nz, cyc_x, cyc_y, A, p, r);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) r[i] = b[i] - r[i];
    const double nr = sqrt(dot64(N, r, r));
    rel = nr / nb;
    if (!isfinite(rel)) {
      status = ORCTM_NH_BREAKDOWN;
      break;
    }
    if (rel <= rtol) {
      status = ORCTM_NH_OK;
      break;
    }
    if (m >= max_outer || inner_total >= maxit) {
      status = ORCTM_NH_NOT_CONVERGED;
      break;
    }
    double tol = 0.5 * rtol * nb / nr;
    if (tol < rtol_inner) tol = rtol_inner;
    const double gamma = 1.0 / (smax * nr_prev), inv_gamma = smax * nr_prev;
    nr_prev = nr;
    int its = 0, ist;
    if (precision == ORCTM_NH_FP64) {
      double* ri = (double*)rin;
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) ri[i] = (S[i] * (-r[i])) * gamma;
      ist = orc_pcg64(nx, ny, nz, cyc_x, cyc_y, (const double*)C, (const double*)f1, (const double*)f2, ri,
                      (double*)x, (double*)work, tol, maxit - inner_total, &its);
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) p[i] = p[i] + ((double*)x)[i] * inv_gamma;
    } else {
      float* ri = (float*)rin;
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) ri[i] = (float)((S[i] * (-r[i])) * gamma);
      ist = orc_pcg32(nx, ny, nz, cyc_x, cyc_y, (const float*)C, (const float*)f1, (const float*)f2, ri, (float*)x,
                      (float*)work, (float)tol, maxit - inner_total, &its);
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) p[i] = p[i] + (double)((float*)x)[i] * inv_gamma;
    }
    inner_total += its;
    if (ist == ORCTM_NH_BREAKDOWN && its == 0) {
      status = ORCTM_NH_BREAKDOWN;
      break;
    }
  }
  s.status = status;
  s.outer_iters = m;
  s.inner_iters = inner_total;
  s.rel_res = rel;
  s.solve_ms = wall_ms() - t_iter;
  s.total_ms = wall_ms() - t_begin;
  if (st) *st = s;
  free(r);
  free(C);
  free(f1);
  free(f2);
  free(rin);
  free(x);
  free(work);
  return status;
}

/* ------------------------------------------------------------------------------------------- */
/* FGMRES(m), right-preconditioned, FP64 -- the reference's KSP type (restart_nonhydro.sh:30).    */
/* Minimises ||b - A x||_2 (unpreconditioned norm); preconditioner z = M^{-1}(S .* (-v)), M the  */
/* line preconditioner of C = S(-A). Stops when the Arnoldi residual estimate <= rtol*||b|| and */
/* the true residual confirms it.                                                                */
int orc_fgmres(int nx, int ny, int nz, int cyc_x, int cyc_y, const double* A, const double* S, const double* b,
               double* x, int restart, double rtol, int maxit, orctm_nh_stats_t* st) {
  const double t_begin = wall_ms();
  const size_t N = (size_t)nx * ny * nz;
  const int mr = restart > 0 ? restart : 30;
  orctm_nh_stats_t s;
  memset(&s, 0, sizeof s);
  const double nb = sqrt(dot64(N, b, b));
  s.bnorm = nb;
  if (!isfinite(nb)) {
    s.status = ORCTM_NH_ERR_NONFINITE;
    if (st) *st = s;
    return ORCTM_NH_ERR_NONFINITE;
  }
  if (nb == 0.0) {
    memset(x, 0, sizeof(double) * N);
    if (st) *st = s;
    return ORCTM_NH_OK;
  }
  double* C = (double*)malloc(sizeof(double) * NC * N);
  double* f1 = (double*)malloc(sizeof(double) * N);
  double* f2 = (double*)malloc(sizeof(double) * N);
  orc_inner_coef64(N, A, S, C);
  orc_line_factor64(nx, ny, nz, C, f1, f2);
  double* V = (double*)malloc(sizeof(double) * N * (size_t)(mr + 1));
  double* Z = (double*)malloc(sizeof(double) * N * (size_t)mr);
  double* w = (double*)malloc(sizeof(double) * N);
  double* t = (double*)malloc(sizeof(double) * N);
  double* H = (double*)calloc((size_t)(mr + 1) * mr, sizeof(double));
  double *cs = (double*)malloc(sizeof(double) * mr), *sn = (double*)malloc(sizeof(double) * mr);
  double* gv = (double*)malloc(sizeof(double) * (mr + 1));
  double* y = (double*)malloc(sizeof(double) * mr);
  int its = 0, status = ORCTM_NH_NOT_CONVERGED;
  double rel = 0.0;
  const double t_iter = wall_ms(); /* set-up (inner operator, line factors, Krylov basis) done */
  while (1) {
    orc_matvec64(nx, ny, nz, cyc_x, cyc_y, A, x, w);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) w[i] = b[i] - w[i];
    const double beta = sqrt(dot64(N, w, w));
    rel = beta / nb;
    if (!isfinite(rel)) {
      status = ORCTM_NH_BREAKDOWN;
      break;
    }
    if (rel <= rtol) {
      status = ORCTM_NH_OK;
      break;
    }
    if (its >= maxit) {
      status = ORCTM_NH_NOT_CONVERGED;
      break;
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) V[i] = w[i] / beta;
    memset(gv, 0, sizeof(double) * (mr + 1));
    gv[0] = beta;
    int j = 0;
    for (; j < mr && its < maxit; ++j, ++its) {
      double* vj = V + (size_t)j * N;
      double* zj = Z + (size_t)j * N;
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) t[i] = S[i] * (-vj[i]);
      orc_line_apply64(nx, ny, nz, f1, f2, t, zj, NULL);
      orc_matvec64(nx, ny, nz, cyc_x, cyc_y, A, zj, w);
      for (int i = 0; i <= j; ++i) { /* modified Gram-Schmidt */
        const double* vi = V + (size_t)i * N;
        const double h = dot64(N, w, vi);
        H[(size_t)i * mr + j] = h;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < N; ++q) w[q] -= h * vi[q];
      }
      const double hn = sqrt(dot64(N, w, w));
      H[(size_t)(j + 1) * mr + j] = hn;
      double* vn = V + (size_t)(j + 1) * N;
      if (hn != 0.0) {
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < N; ++q) vn[q] = w[q] / hn;
      }
      for (int i = 0; i < j; ++i) { /* previous Givens rotations */
        const double a = H[(size_t)i * mr + j], c = H[(size_t)(i + 1) * mr + j];
        H[(size_t)i * mr + j] = cs[i] * a + sn[i] * c;
        H[(size_t)(i + 1) * mr + j] = -sn[i] * a + cs[i] * c;
      }
      const double a = H[(size_t)j * mr + j], c = H[(size_t)(j + 1) * mr + j];
      const double den = hypot(a, c);
      cs[j] = den == 0.0 ? 1.0 : a / den;
      sn[j] = den == 0.0 ? 0.0 : c / den;
      H[(size_t)j * mr + j] = den;
      H[(size_t)(j + 1) * mr + j] = 0.0;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      if (fabs(gv[j + 1]) <= rtol * nb || hn == 0.0) {
        ++j;
        ++its;
        break;
      }
    }
    for (int i = j - 1; i >= 0; --i) { /* back substitution */
      double sum = gv[i];
      for (int q = i + 1; q < j; ++q) sum -= H[(size_t)i * mr + q] * y[q];
      y[i] = sum / H[(size_t)i * mr + i];
    }
    for (int i = 0; i < j; ++i) {
      const double* zi = Z + (size_t)i * N;
      const double yi = y[i];
#pragma omp parallel for schedule(static)
      for (size_t q = 0; q < N; ++q) x[q] += yi * zi[q];
    }
  }
  s.status = status;
  s.outer_iters = 0;
  s.inner_iters = its;
  s.rel_res = rel;
  s.solve_ms = wall_ms() - t_iter;
  s.total_ms = wall_ms() - t_begin;
  if (st) *st = s;
  free(C);
  free(f1);
  free(f2);
  free(V);
  free(Z);
  free(w);
  free(t);
  free(H);
  free(cs);
  free(sn);
  free(gv);
  free(y);
  return status;
}

/* Number of threads OpenMP will use (reported as the CPU baseline's core count). */
#ifdef _OPENMP
#include <omp.h>
int orc_num_threads(void) { return omp_get_max_threads(); }
#else
int orc_num_threads(void) { return 1; }
#endif
