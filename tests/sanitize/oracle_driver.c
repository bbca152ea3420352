/*
 * Sanitizer driver of the CPU oracle (tests/test_sanitizers.py): builds small seeded grids of the shape
 * ORCTM hands over (Fortran arrays with halos: land, partial bottom cells, open, closed and periodic
 * sides) and calls every oracle entry point on them, so that an AddressSanitizer / UndefinedBehavior
 * Sanitizer build of oracle/ reports any out-of-bounds access, use of freed memory or undefined
 * arithmetic. Prints "SANITIZED OK" when every call returned and the results are finite.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/orctm_nh.h"
#include "../../oracle/nh_oracle_bot.h"
#include "../../oracle/nh_oracle_ocv.h"
#include "../../oracle/nh_oracle_sor.h"

int orc_assemble(const orctm_nh_grid_t* g, double* A, double* S);
int orc_compact(const orctm_nh_grid_t* g, double* zw, double* zp, uint16_t* flags, double* col);
void orc_rebuild64(int nx, int ny, int nz, const double* zw, const double* zp, const uint16_t* flags,
                   const double* col, double* A);
int orc_rhs(const orctm_nh_grid_t* g, const double* divg, double* b);
int orc_scatter_interior(const orctm_nh_grid_t* g, const double* x, double* pnh);
int orc_solve(int nx, int ny, int nz, int cyc_x, int cyc_y, const double* A, const double* S, const double* b,
              double* p, int precision, double rtol, double rtol_inner, int maxit, int max_outer,
              orctm_nh_stats_t* st, const void* Cin);
int orc_fgmres(int nx, int ny, int nz, int cyc_x, int cyc_y, const double* A, const double* S, const double* b,
               double* x, int restart, double rtol, int maxit, orctm_nh_stats_t* st);

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static double urand(void) { /* SplitMix64 -> [0, 1) */
  uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static double* alloc(size_t n) {
  double* p = (double*)calloc(n ? n : 1, sizeof(double));
  if (!p) abort();
  return p;
}

static int finite_all(const double* a, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!isfinite(a[i])) return 0;
  return 1;
}

/* One case: IE x JE x KE with the given sides; returns the number of failed checks. */
static int run_case(int IE, int JE, int KE, int cx, int cy, int open) {
  const int IS = 0, JS = 0;
  const size_t p2 = (size_t)IE * JE, p3 = p2 * KE, u2 = (size_t)(IE - IS + 1) * JE, v2 = (size_t)IE * (JE - JS + 1);
  const size_t u3 = u2 * KE, v3 = v2 * KE, w3 = p2 * (KE + 1);
  double *dlxp = alloc(p2), *dlyp = alloc(p2), *dlxu = alloc(u2), *dlyv = alloc(v2), *dzw = alloc(KE);
  double *ddpo = alloc(p3), *weto = alloc(p3), *divg = alloc(p3), *pnh = alloc(p3);
  for (int k = 0; k < KE; ++k) dzw[k] = 10.0 + 5.0 * k;
  double depth = 0.0;
  for (int k = 0; k < KE; ++k) depth += dzw[k];
  for (size_t c = 0; c < p2; ++c) {
    dlxp[c] = 900.0 + 400.0 * urand();
    dlyp[c] = 1000.0 + 300.0 * urand();
  }
  for (size_t c = 0; c < u2; ++c) dlxu[c] = 900.0 + 400.0 * urand();
  for (size_t c = 0; c < v2; ++c) dlyv[c] = 1000.0 + 300.0 * urand();
  for (int j = 0; j < JE; ++j)
    for (int i = 0; i < IE; ++i) {
      const double d = urand() < 0.2 ? 0.0 : depth * (0.1 + 0.9 * urand());
      double top = 0.0;
      for (int k = 0; k < KE; ++k) {
        const size_t c = (size_t)i + (size_t)IE * ((size_t)j + (size_t)JE * k);
        const double mid = top + 0.5 * dzw[k];
        weto[c] = mid < d ? 1.0 : 0.0;
        ddpo[c] = weto[c] == 0.0 ? 0.0 : (top + dzw[k] <= d ? dzw[k] : d - top);
        divg[c] = weto[c] * (urand() - 0.5);
        top += dzw[k];
      }
    }
  orctm_nh_grid_t g;
  memset(&g, 0, sizeof g);
  g.ie = IE, g.je = JE, g.ke = KE, g.i_start = IS, g.j_start = JS;
  g.dlxp = dlxp, g.dlyp = dlyp, g.dlxu = dlxu, g.dlyv = dlyv, g.ddpo = ddpo, g.weto = weto, g.dzw = dzw;
  g.cyclic_x = cx, g.cyclic_y = cy, g.open_w = open, g.open_e = open, g.open_n = open && !cy, g.open_s = 0;
  const int nx = IE - 2, ny = JE - 2, nz = KE;
  const size_t N = (size_t)nx * ny * nz, P = (size_t)nx * ny;
  int fails = 0;
  double *A = alloc(7 * N), *S = alloc(N), *b = alloc(N), *p = alloc(N), *x = alloc(N), *A2 = alloc(7 * N);
  double *zw = alloc((size_t)(nz + 1) * P), *zp = alloc(N), *col = alloc(7 * P);
  uint16_t* fl = (uint16_t*)calloc(N, sizeof(uint16_t));
  fails += orc_assemble(&g, A, S) != ORCTM_NH_OK;
  fails += orc_compact(&g, zw, zp, fl, col) != ORCTM_NH_OK;
  orc_rebuild64(nx, ny, nz, zw, zp, fl, col, A2);
  fails += memcmp(A, A2, sizeof(double) * 7 * N) != 0;
  fails += orc_rhs(&g, divg, b) != ORCTM_NH_OK;
  orctm_nh_stats_t st;
  for (int prec = 0; prec <= 1; ++prec) {
    memset(p, 0, sizeof(double) * N);
    fails += orc_solve(nx, ny, nz, cx, cy, A, S, b, p, prec, 1e-10, -1.0, 10000, 40, &st, NULL) != ORCTM_NH_OK;
    fails += !finite_all(p, N);
  }
  memset(x, 0, sizeof(double) * N);
  fails += orc_fgmres(nx, ny, nz, cx, cy, A, S, b, x, 30, 1e-10, 20000, &st) != ORCTM_NH_OK;
  fails += orc_scatter_interior(&g, p, pnh) != ORCTM_NH_OK;

  /* OCVTOT_NH around the solve and the bottom-pressure solve */
  double *amsuo = alloc(u3), *dduo = alloc(u3), *dtdxuo = alloc(u2), *dlyu = alloc(u2);
  double *amsue = alloc(v3), *ddue = alloc(v3), *dpye = alloc(v2), *dlxv = alloc(v2), *dtdxpo = alloc(p2),
         *dtdyo = alloc(p2);
  double *uoo = alloc(u3), *uoo1 = alloc(u3), *voe = alloc(v3), *voe1 = alloc(v3), *wno = alloc(w3),
         *wno1 = alloc(w3), *zo = alloc(p2);
  for (size_t c = 0; c < u3; ++c) {
    amsuo[c] = urand() < 0.8 ? 1.0 : 0.0;
    dduo[c] = amsuo[c] * 12.0;
    uoo[c] = amsuo[c] * (urand() - 0.5);
    uoo1[c] = amsuo[c] * (urand() - 0.5);
  }
  for (size_t c = 0; c < v3; ++c) {
    amsue[c] = urand() < 0.8 ? 1.0 : 0.0;
    ddue[c] = amsue[c] * 12.0;
    voe[c] = amsue[c] * (urand() - 0.5);
    voe1[c] = amsue[c] * (urand() - 0.5);
  }
  for (size_t c = 0; c < u2; ++c) dtdxuo[c] = 60.0 / dlxu[c], dlyu[c] = 1100.0;
  for (size_t c = 0; c < v2; ++c) dpye[c] = 60.0 / dlyv[c], dlxv[c] = 1000.0;
  for (size_t c = 0; c < p2; ++c) dtdxpo[c] = 60.0 / dlxp[c], dtdyo[c] = 60.0 / dlyp[c], zo[c] = 0.01 * urand();
  for (size_t c = 0; c < w3; ++c) wno[c] = 1e-3 * (urand() - 0.5), wno1[c] = 1e-3 * (urand() - 0.5);
  orc_vgrid vg;
  memset(&vg, 0, sizeof vg);
  vg.IE = IE, vg.JE = JE, vg.KE = KE, vg.i_start = IS, vg.j_start = JS, vg.cyc_x = cx, vg.cyc_y = cy;
  vg.open_w = g.open_w, vg.open_e = g.open_e, vg.open_n = g.open_n, vg.open_s = g.open_s;
  vg.dlxp = dlxp, vg.dlyp = dlyp, vg.ddpo = ddpo, vg.weto = weto, vg.amsuo = amsuo, vg.dduo = dduo;
  vg.dtdxuo = dtdxuo, vg.dlyu = dlyu, vg.amsue = amsue, vg.ddue = ddue, vg.dpye = dpye, vg.dlxv = dlxv;
  vg.dtdxpo = dtdxpo, vg.dtdyo = dtdyo, vg.dt = 60.0, vg.dti = 1.0 / 60.0;
  double* d2 = alloc(p3);
  orc_divergence(&vg, uoo, uoo1, voe, voe1, wno, wno1, zo, d2);
  orc_project(&vg, pnh, uoo, voe, wno);
  fails += !finite_all(d2, p3) + !finite_all(uoo, u3) + !finite_all(voe, v3) + !finite_all(wno, w3);
  double *bu = alloc(u2), *bv = alloc(v2), *bw = alloc(p2), *dbs = alloc(p2), *coef = alloc(5 * (size_t)nx * ny);
  double *b2 = alloc((size_t)nx * ny), *x2 = alloc((size_t)nx * ny), *z2 = alloc((size_t)nx * ny), *pb = alloc(p2);
  orc_bottom_rhs(&vg, uoo, uoo1, voe, voe1, wno, wno1, d2, bu, bv, bw, dbs);
  orc_op2d_coef(&g, coef);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) b2[(size_t)j * nx + i] = dbs[(size_t)(i + 1) + (size_t)IE * (j + 1)];
  double rr = 0.0;
  orc_solve2d(&g, coef, b2, x2, 1e-10, 2000, &rr);
  fails += !(rr <= 1e-10) || !finite_all(x2, (size_t)nx * ny);
  orc_mg_apply(&g, coef, b2, z2);
  orc_op2d_apply(&g, coef, x2, z2);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) pb[(size_t)(i + 1) + (size_t)IE * (j + 1)] = x2[(size_t)j * nx + i];
  orc_bounds_exch(pb, IE, JE, 1, 1, 1, cx, cy);
  orc_bottom_correction(&g, pb, d2);
  fails += !finite_all(d2, p3);

  /* ITPREP / TRONEU */
  double *xx = alloc(p2), *uf = alloc(u2), *vf = alloc(v2), *ff = alloc(p2), *w1 = alloc(p2);
  double *deuto = alloc(u2), *deute = alloc(v2), *u1o = alloc(u2), *uzo = alloc(u2), *v1e = alloc(v2),
         *vze = alloc(v2), *b1o = alloc(p2), *z1o = alloc(p2);
  for (size_t c = 0; c < p2; ++c) w1[c] = weto[c];
  for (size_t c = 0; c < u2; ++c) deuto[c] = 500.0 * urand(), u1o[c] = urand() - 0.5, uzo[c] = urand() - 0.5;
  for (size_t c = 0; c < v2; ++c) deute[c] = 500.0 * urand(), v1e[c] = urand() - 0.5, vze[c] = urand() - 0.5;
  orc_sorgrid sg;
  memset(&sg, 0, sizeof sg);
  sg.IE = IE, sg.JE = JE, sg.i_start = IS, sg.j_start = JS, sg.cyc_x = cx, sg.cyc_y = cy, sg.iterations = 20;
  sg.dlxp = dlxp, sg.dlyp = dlyp, sg.dtdxpo = dtdxpo, sg.dtdyo = dtdyo, sg.weto1 = w1;
  sg.dlxu = dlxu, sg.dlyu = dlyu, sg.deuto = deuto, sg.dlxv = dlxv, sg.dlyv = dlyv, sg.deute = deute;
  sg.g = 9.81, sg.conn = 0.5, sg.cono = 0.5, sg.stabn = 0.5, sg.dt = 60.0, sg.sorpar = 1.7;
  orc_itprep(&sg, xx, uf, vf, ff);
  orc_troneu(&sg, xx, uf, vf, ff, u1o, uzo, v1e, vze, b1o, z1o);
  fails += !finite_all(z1o, p2);

  void* all[] = {dlxp, dlyp, dlxu, dlyv, dzw, ddpo, weto, divg, pnh, A, S, b, p, x, A2, zw, zp, col, fl,
                 amsuo, dduo, dtdxuo, dlyu, amsue, ddue, dpye, dlxv, dtdxpo, dtdyo, uoo, uoo1, voe, voe1, wno,
                 wno1, zo, d2, bu, bv, bw, dbs, coef, b2, x2, z2, pb, xx, uf, vf, ff, w1, deuto, deute, u1o, uzo,
                 v1e, vze, b1o, z1o};
  for (size_t q = 0; q < sizeof all / sizeof all[0]; ++q) free(all[q]);
  printf("case %dx%dx%d cyc %d/%d open %d: %s\n", IE, JE, KE, cx, cy, open, fails ? "FAIL" : "ok");
  return fails;
}

int main(void) {
  int fails = 0;
  fails += run_case(19, 13, 7, 0, 0, 1);
  fails += run_case(16, 11, 5, 1, 0, 0);
  fails += run_case(14, 12, 6, 0, 1, 1);
  fails += run_case(12, 10, 1, 1, 1, 0);
  fails += run_case(5, 4, 3, 0, 0, 0);
  printf("%s\n", fails ? "SANITIZED FAILURES" : "SANITIZED OK");
  return fails != 0;
}
