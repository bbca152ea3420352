// nh_kernels.h -- host-side launchers of the gfx950 Poisson kernels (nh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "nh_device.h"

namespace onh {

// Device copies of the caller's Fortran arrays for one assembly (mo_petsc.f90:543-782 inputs).
struct FortranIn {
  int IE, JE, KE, i_start, j_start;
  int cyc_x, cyc_y, open_w, open_e, open_n, open_s;
  const double *dlxp, *dlyp, *dlxu, *dlyv, *ddpo, *weto, *dzw;
};

// Assembly: A64, S, C (inner precision), wet; flags invalid input in *bad; max S in *smax_bits.
hipError_t launch_assemble(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                           const FortranIn& f, int* d_bad, unsigned long long* d_smax_bits);
hipError_t launch_line_factor(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles);
// b = DIVG*WETO from a device copy of DIVG(IE,JE,KE); ||b|| into Outer.nb.
hipError_t launch_rhs(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* d_divg, int IE,
                      int JE);
// Interior of a device Fortran (IE,JE,KE) array -> p64 (initial guess).
hipError_t launch_guess(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* d_pnh, int IE,
                        int JE);
// Per-system outer-state initialisation (nsys threads).
hipError_t launch_begin(hipStream_t st, bool fp32, const Sys* d_sys, int nsys);
hipError_t launch_residual(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles);
hipError_t launch_init(hipStream_t st, bool fp32, int nzr, const Sys* d_sys, const Tile* d_tiles, int ntiles);
hipError_t launch_spmv(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles);
hipError_t launch_update(hipStream_t st, bool fp32, int nzr, const Sys* d_sys, const Tile* d_tiles, int ntiles);
hipError_t launch_refine(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles);
// y = A64 x on padded interior vectors (verification entry point).
hipError_t launch_apply64(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* x,
                          double* y);

// Register-array length used by the column kernels for a given nz (0 = global scratch path).
int pick_nzr(bool fp32, int nz);

}  // namespace onh
