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

// Assembly into the compact form (ZW, ZP, flags, column products and their inner-precision copies);
// flags invalid input in *bad; max S in *smax_bits.
hipError_t launch_assemble(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                           const FortranIn& f, int* d_bad, unsigned long long* d_smax_bits);
hipError_t launch_line_factor(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles);
// The per-level ZW/ZP table of the wet column and each column's 3-level window (Sys::tab); *d_bad set
// when some column's deviations from the table do not fit in 3 consecutive levels.
hipError_t launch_zwindow(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                          const FortranIn& f, int* d_bad);
// b = DIVG*WETO from a device copy of DIVG(IE,JE,KE); ||b|| into Outer.nb.
hipError_t launch_rhs(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                      const double* d_divg, int IE, int JE);
// b <- b (1 + amplitude xi), xi ~ N(0,1) from (d_seeds[system], cell); ||b|| refreshed.
hipError_t launch_perturb(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                          double amplitude, const unsigned long long* d_seeds);
// Interior of a device Fortran (IE,JE,KE) array -> p64 (initial guess).
hipError_t launch_guess(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* d_pnh, int IE,
                        int JE);
// Per-system outer-state initialisation (nsys threads).
hipError_t launch_begin(hipStream_t st, bool fp32, const Sys* d_sys, int nsys);
// Rows per tile of the residual kernel: 64 x kOuterRows blocks. It holds 112 registers per thread (4 waves
// per SIMD): blocks of 4 wavefronts fill 16 waves per CU where 12-row blocks (12 wavefronts) leave room for
// only one (1024x768x60: 308 -> 285 us FP32, 382 -> 365 us FP64). The update kernel k_refine (44
// registers, 8 waves per SIMD) runs on the set-up tiling, where it is faster (194 against 203 us).
// h = kOuterRows or TY (the set-up tiling).
constexpr int kOuterRows = 4;
// Each phase launch below is followed by its one-block-per-system reduction (nsys blocks).
hipError_t launch_residual(hipStream_t st, bool fp32, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                           int nsys);
// Inner iteration kernels on tiles of 64 x h columns (h in kTileHeights; d_tiles built for that h).
// Each launch over one group of tiles; the per-system reduction of the phase follows once all
// groups of the phase have been launched. init: x = 0 and the first forward elimination.
constexpr int kTileHeights[3] = {4, 8, 12};
enum InnerPhase { PH_INIT = 0, PH_BWD = 1, PH_FWD = 2 };
// A launch of fewer tiles than this is a few chains of dependent levels, and the kernels take their
// deep-prefetch or chunked forms. Tile lists of streaming launches (this many real tiles or more) may
// hold empty tiles (sys < 0), which the streaming kernels skip; the count passed is the list's length.
constexpr int kDeepTiles = 256;
// d_pairs: the tiles of d_tiles with an even bx (one block runs that tile and the next one in x), for
// launches of kDeepTiles tiles or more.
hipError_t launch_fwd(hipStream_t st, bool fp32, bool init, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                      const Tile* d_pairs, int npairs);
// raw = 1 (verification only): q = C w without the back substitution. buf: the kernel may address
// memory through buffer descriptors, which requires every inner array of every system of the launch
// to hold at most kBufMaxBytes (bwd_buffers_fit); else 64-bit pointers. Same bits either way.
hipError_t launch_bwd(hipStream_t st, bool fp32, bool buf, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                      int raw);
// The largest array the buffer-addressed matvec handles (lane offsets stay below 2^31 bytes).
constexpr unsigned long long kBufMaxBytes = 1ull << 31;
// Whether a system's inner arrays (nz planes in the inner precision) fit the buffer-addressed matvec.
inline bool bwd_buffers_fit(const Sys& s, bool fp32) {
  return (unsigned long long)s.plane * (unsigned long long)s.nz * (fp32 ? 4ull : 8ull) <= kBufMaxBytes;
}
hipError_t launch_inner_reduce(hipStream_t st, bool fp32, InnerPhase ph, const Sys* d_sys, int nsys);
hipError_t launch_refine(hipStream_t st, bool fp32, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles);
// The 7 planes of the reference operator rebuilt from the compact form (verification entry point).
hipError_t launch_materialize(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, double* A);
// dst <- src, bytes rounded down to 16 (bandwidth probe).
hipError_t launch_copy4(hipStream_t st, const void* src, void* dst, size_t bytes);
// y = A64 x on padded interior vectors (verification entry point).
hipError_t launch_apply64(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* x,
                          double* y);

}  // namespace onh
