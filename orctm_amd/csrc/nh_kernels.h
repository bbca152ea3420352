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
hipError_t launch_rhs(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                      const double* d_divg, int IE, int JE);
// b <- b (1 + amplitude xi), xi ~ N(0,1) from (seed, cell); ||b|| refreshed.
hipError_t launch_perturb(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                          double amplitude, unsigned long long seed);
// Interior of a device Fortran (IE,JE,KE) array -> p64 (initial guess).
hipError_t launch_guess(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* d_pnh, int IE,
                        int JE);
// Per-system outer-state initialisation (nsys threads).
hipError_t launch_begin(hipStream_t st, bool fp32, const Sys* d_sys, int nsys);
// Each phase launch below is followed by its one-block-per-system reduction (nsys blocks).
hipError_t launch_residual(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys);
hipError_t launch_init(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys);
// raw = 1 (verification only): q = C w without the back substitution.
hipError_t launch_spmv(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                       int raw = 0);
hipError_t launch_update(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys);
hipError_t launch_refine(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles);
// dst <- src, bytes rounded down to 16 (bandwidth probe).
hipError_t launch_copy4(hipStream_t st, const void* src, void* dst, size_t bytes);
// y = A64 x on padded interior vectors (verification entry point).
hipError_t launch_apply64(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* x,
                          double* y);

}  // namespace onh
