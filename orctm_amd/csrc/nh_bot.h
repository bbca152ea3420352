// nh_bot.h -- the 2-D bottom-pressure solve of the CG2D4NONHY option on the device (nh_bot.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "nh_ocv.h"

namespace onh {

// Device state of the bottom solve of one handle. Arrays in the caller's Fortran layouts with halos,
// except the box vectors, which hold the (IE-2) x (JE-2) interior box in DMDA order (i fastest).
struct BotState {
  int IE = 0, JE = 0, KE = 0, IS = 1, JS = 1, cyc_x = 0, cyc_y = 0, nx = 0, ny = 0;
  const double *dlxu = nullptr, *dlyv = nullptr, *dlxp = nullptr, *dlyp = nullptr, *dzw = nullptr;
  double* coef = nullptr;  // 5 planes of nx*ny: centre, west, east, north, south of compute_matrix_2d
  double* S = nullptr;     // DLXP*DLYP per box cell
  double *botgu = nullptr, *botgv = nullptr, *botgw = nullptr;  // module arrays kept between steps
  double *dbs = nullptr, *pnh_bot = nullptr;                      // DIVG_bot_surf, PNH_bot (IE,JE)
  double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *d0 = nullptr, *d1 = nullptr, *q = nullptr;
  double* part = nullptr;  // block partial sums, 3 per block
  // Multigrid preconditioner of the 2-D PCG, built by bot_setup: level 0 is the box (its edge rows are
  // not unknowns), each next level aggregates 2 x 2 cells of the one before with the Galerkin operator
  // (again 5-point), and the last level (at most kMgDense unknowns) is solved with a dense inverse.
  struct MgLevel {
    int a = 0, b = 0, o = 0;  // dimensions; the unknowns are i in [o, a-o), j in [o, b-o)
    double* k = nullptr;      // 5 planes of a*b: centre, west, east, north (j-1), south (j+1)
    double *x = nullptr, *y = nullptr, *r = nullptr;  // iterate pair and right-hand side
  };
  std::vector<MgLevel> mg;
  double* mg_inv = nullptr;  // inverse of the last level's operator over its unknowns, row-major
  int mg_n = 0;
  std::vector<void*> mg_allocs;
  double* scal = nullptr;  // rz, dq, alpha, beta, rr, bb, breakdown
  double* scal_host = nullptr;  // pinned copy of scal
  int nblk = 0;
  hipGraphExec_t chunk = nullptr;  // kBotChunk iterations
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the last bot_step
  int iters = -1;                  // iterations of the last solve, -1 before the first
  double rel_res = 0.0;
  int status = 0;                  // of the last solve; a solve after one that was not OK starts cold
  double sym_defect = 0.0;         // ||K - K^T||_F / ||K||_F of the box operator K = -S A2 (bot_setup)
};

constexpr int kBotChunk = 8;   // PCG iterations per captured graph, between residual checks
constexpr int kMgDense = 64;   // unknowns of the last multigrid level at most

// coef and S from the metrics (mo_petsc.f90:401-538) and the multigrid hierarchy (allocated here,
// freed by bot_release). The caller allocates every other array of BotState (zero-filled: the module
// arrays start at 0 as beleg_zero.f90:481-504 sets them), scal_host pinned and the two events.
hipError_t bot_setup(hipStream_t st, BotState& b);
// ocvtot_nh.f90:64-147 and petsc_solve_2D (mo_petsc.f90:147-166): BOTGU/BOTGV/BOTGW, DIVG_bot_surf from
// the velocity state and DIVG (first level, interior) of this step, then PNH_bot from the 2-D solve
// (warm start from the previous solution) and its bounds_exch. b.status, b.iters, b.rel_res report it.
hipError_t bot_step(hipStream_t st, BotState& b, const OcvArrays& a, const double* divg, double rtol, int maxit);
// ComputeRHS with CG2D4NONHY (mo_petsc.f90:252-374): DIVG corrected in place on the interior.
hipError_t bot_correct(hipStream_t st, const BotState& b, const OcvArrays& a, double* divg);
void bot_release(BotState& b);

}  // namespace onh
