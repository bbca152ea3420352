/* nh_oracle_ocv.h -- CPU restatement of OCVTOT_NH around the solve (TEST INFRASTRUCTURE ONLY). */
#pragma once

/* Sizes, flags, metrics and step constants; the arrays are the caller's Fortran arrays with halos. */
typedef struct orc_vgrid {
  int IE, JE, KE, i_start, j_start;
  int cyc_x, cyc_y, open_w, open_e, open_n, open_s;
  const double *dlxp, *dlyp, *ddpo, *weto;    /* (IE,JE) x2, (IE,JE,KE) x2 */
  const double *amsuo, *dduo, *dtdxuo, *dlyu; /* (I_start:IE,JE,KE) x2, (I_start:IE,JE) x2 */
  const double *amsue, *ddue, *dpye, *dlxv;   /* (IE,J_start:JE,KE) x2, (IE,J_start:JE) x2 */
  const double *dtdxpo, *dtdyo;               /* (IE,JE) */
  double dt, dti;
} orc_vgrid;

void orc_bounds_exch(double* a, int IE, int JE, int nk, int i0, int j0, int cyc_x, int cyc_y);
void orc_divergence(const orc_vgrid* g, const double* uoo, const double* uoo_1, const double* voe,
                    const double* voe_1, const double* wno, const double* wno_1, const double* zo, double* divg);
void orc_project(const orc_vgrid* g, double* pnh, double* uoo, double* voe, double* wno);
