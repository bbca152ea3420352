/* nh_oracle_bot.h -- CPU restatement of the 2-D bottom-pressure solve (CG2D4NONHY) around the 3-D
 * nonhydrostatic solve (TEST INFRASTRUCTURE ONLY). */
#pragma once

#include "../include/orctm_nh.h"
#include "nh_oracle_ocv.h"

/* ocvtot_nh.f90:64-144: BOTGU (I_start:IE,JE), BOTGV (IE,J_start:JE), BOTGW (IE,JE) are read and written
 * (the module arrays keep their values between steps where the source assigns nothing); divg is DIVG of
 * the same step ((IE,JE,KE), its first level read on the interior); dbs receives DIVG_bot_surf on
 * i = 2..IE-1, j = 2..JE-1 (the cells ComputeRHS_2d reads, mo_petsc.f90:178-182), its halo untouched. */
void orc_bottom_rhs(const orc_vgrid* g, const double* uoo, const double* uoo_1, const double* voe,
                    const double* voe_1, const double* wno, const double* wno_1, const double* divg, double* botgu,
                    double* botgv, double* botgw, double* dbs);

/* compute_matrix_2d (mo_petsc.f90:401-538) with Neumann = .FALSE. (:37), on the interior box
 * nx = IE-2 by ny = JE-2 (DMDA order, i fastest): coef receives 5 planes of nx*ny values, 0 centre,
 * 1 west (i-1), 2 east (i+1), 3 north (j-1), 4 south (j+1); rows on the edge of the box keep only the
 * centre -(X_C+Y_C). */
void orc_op2d_coef(const orctm_nh_grid_t* g, double* coef);

/* y = A2 p on the interior box, each row summed in column order (north, west, centre, east, south) as
 * an AIJ MatMult does. */
void orc_op2d_apply(const orctm_nh_grid_t* g, const double* coef, const double* p, double* y);

/* The 2-D solve as the device runs it, in FP64 on the CPU (no counterpart in the reference, whose cg2d_
 * KSP runs inside PETSc): the box-edge rows directly (p = b / A_cc), then CG on the remaining rows scaled
 * by S = DLXP*DLYP, preconditioned by one multigrid V-cycle (2 x 2 aggregates, Galerkin operators, 2
 * damped-Jacobi sweeps before and after each coarse correction, corrections scaled by 1.8, a dense solve
 * on the last level of at most 64 unknowns), from x (the warm start, edge values ignored), checking
 * ||b - A2 p||_2 <= rtol ||b||_2 every 8 iterations. Returns the iterations; *rel_res the final ratio. */
int orc_solve2d(const orctm_nh_grid_t* g, const double* coef, const double* b, double* x, double rtol, int maxit,
                double* rel_res);

/* One V-cycle of that preconditioner, z = M r, on the box (r on the box edge ignored, z 0 there).
 * Returns the number of levels (0 if the box has no inner rows). */
int orc_mg_apply(const orctm_nh_grid_t* g, const double* coef, const double* r, double* z);

/* ComputeRHS with CG2D4NONHY (mo_petsc.f90:252-374): the corrections of DIVG on i = 2..IE-1,
 * j = 2..JE-1, k = 1..KE by PNH_bot (IE,JE, halo included) at dry neighbours and the bottom face, in place. */
void orc_bottom_correction(const orctm_nh_grid_t* g, const double* pnh_bot, double* divg);
