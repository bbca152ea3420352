/* nh_oracle_sor.h -- CPU restatement of ITPREP / TRONEU, the barotropic SOR solve (TEST INFRASTRUCTURE ONLY). */
#pragma once

typedef struct orc_sorgrid {
  int IE, JE, i_start, j_start, cyc_x, cyc_y, iterations;
  const double *dlxp, *dlyp, *dtdxpo, *dtdyo, *weto1; /* (IE,JE) */
  const double *dlxu, *dlyu, *deuto;                  /* (I_start:IE,JE) */
  const double *dlxv, *dlyv, *deute;                  /* (IE,J_start:JE) */
  double g, conn, cono, stabn, dt, sorpar;
} orc_sorgrid;

void orc_itprep(const orc_sorgrid* g, double* xx, double* uf, double* vf, double* ff);
void orc_troneu(const orc_sorgrid* g, const double* xx, const double* uf, const double* vf, const double* ff,
                const double* u1o, const double* uzo, const double* v1e, const double* vze, double* b1o,
                double* z1o);
