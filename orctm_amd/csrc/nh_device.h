// nh_device.h -- device-side data structures of the gfx950 Poisson solver.
//
// Layout in HBM (one system): the interior nx x ny x nz cells of the reference DMDA
// (mo_petsc.f90:48-50,75-95) stored i-fastest with a row pitch padded to a multiple of 32
// elements: cell (i,j,k) at k*plane + j*pitch + i, plane = pitch*ny. The operator is kept in a
// compact form (ZW, ZP per cell, 12 flag bits per cell, 7 metric products per column) from which
// every row of compute_matrix is rebuilt on the fly with the assembly's own operations. Padding
// cells are never read as data: every neighbour access goes through the wrap/validity map.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace onh {

constexpr int TX = 64;           // tile width in i: one wavefront, consecutive columns
#ifndef ORCTM_TILE_Y
#define ORCTM_TILE_Y 12
#endif
constexpr int TY = ORCTM_TILE_Y; // tile height in j of the set-up kernels: one wavefront per row
constexpr int RG = 4;            // rows per reduction group (block partial sums are per 4 rows)
static_assert(TY % RG == 0 && TY <= 16, "tiles hold whole reduction groups; blocks hold <= 1024 threads");
constexpr int NT = TX * TY;      // threads per block of the set-up kernels
constexpr int NCOEF = 7;

enum Done : int { RUN = 0, CONVERGED = 1, NOTCONV = 2, BREAKDOWN = 3, NONFINITE = 4, SKIP = 5 };

// FP64 outer-refinement state of one system.
struct Outer {
  double nb;          // ||b||_2
  double nr;          // ||b - A p||_2 of the last residual
  double nr_prev;     // normaliser of the current inner right-hand side
  double smax;        // max S over the system (S = DLXP*DLYP/ZP)
  double rtol, rtol_inner;
  double inv_gamma;   // scale of the pending inner solution (p += x * inv_gamma)
  double rel;         // nr / nb
  int outer, max_outer, inner_total, maxit;
  int odone;          // Done
  int pad[3];
};

// Inner PCG state of one system (scalars in the inner precision T).
template <typename T>
struct Inner {
  T rho, pq, rr, rr0, alpha, beta, tol2, padT;
  int iter, maxit, done, first, pbuf;
  int xpend;  // 1: x += alpha p of the last iteration still to be applied
  int pad[2];
};

// One system as the kernels see it (a copy lives in device memory).
struct Sys {
  int nx, ny, nz, pitch;
  long long plane;
  int cyc_x, cyc_y;
  int tiles_x, tiles_y, ntiles;  // set-up tiling (64 x TY)
  int nslots;                    // partial-sum slots: tiles_x * ceil(ny / 4), one per 64 x 4 group
  // operator in compact form (rows are rebuilt on the fly, see row_coef in nh_kernels.hip)
  const double* zw;     // ZW, nz+1 planes (mo_petsc.f90:557-583)
  const double* zp;     // ZP, nz planes (mo_petsc.f90:585-590)
  const uint16_t* flg;  // per-cell flag bits: wet, couplings present, X_C/Y_C/Z_C variants
  const double* col;    // 7 per-column planes: XUw*XP, XUe*XP, YVn*YP, YVs*YP, X_C, Y_C, DLXP*DLYP
  const void* zw_in;    // the same in the inner precision (FP32 copies for MIXED; aliases for FP64)
  const void* zp_in;
  const void* col_in;
  // ZW/ZP as a per-level table plus a 3-level window per column: on a z-level grid they equal the
  // values of a column wet down to KE with full cells everywhere except around the column's bottom
  // cell. The set-up sets `tab` = 1 when it found every column of the system so (bit for bit, both
  // precisions); a solve context keeps it or clears it (options.zw_table), and with it set the
  // iteration kernels read the table and the window instead of the per-cell planes.
  int tab;
  const double* tw;     // nz+1 ZW of the wet column
  const double* tp;     // nz ZP of the wet column
  const void* tw_in;    // the same in the inner precision
  const void* tp_in;
  const int* xk;        // per column: first level of the window
  const double* xw;     // 3 planes: ZW at levels xk, xk+1, xk+2
  const double* xp;     // 3 planes: ZP at levels xk, xk+1, xk+2
  const void* xw_in;    // the same in the inner precision
  const void* xp_in;
  // per column, 3 words: the wet depth (leading wet levels of WETO) of the column and of its W, E, N, S
  // neighbours (halo included), 16 bits each, and 4 bits saying whether the W/E/N/S coupling can exist
  // (periodic or not on the domain edge); with `tab` set the kernels derive each cell's flag bits from
  // them instead of reading the flag plane (checked bit for bit against it at create)
  const uint32_t* xf;
  const void* invD;     // line preconditioner 1/D_k: FP64, or fp16 bits scaled by dscale (FP32)
  const void* L;        // line preconditioner l_k: FP64, or fp16 bits (FP32)
  const float* dscale;  // FP32 only: per-column power of two that multiplies the stored invD
  double* b;            // right-hand side DIVG*WETO
  double* p64;          // solution (FP64)
  void* x;              // inner solution
  void* r;              // inner residual (also the inner right-hand side at init)
  void* z;              // preconditioned residual
  void* pa;             // search direction, buffer A
  void* pb;             // search direction, buffer B
  void* q;              // C p
  double* partials;     // 2*nslots partial sums (summed by the reduction kernel in slot order)
  Outer* outer;
  void* inner;          // Inner<T>*
  double* nbp;          // ||b||_2 of the resident right-hand side (written by the rhs kernel)
  const double* smaxp;  // max S of the system (written by the assembly)
  double rtol, rtol_inner;
  int maxit, max_outer;
};

struct Tile {
  int sys;    // index into the Sys array of the launch
  int bx, by; // tile coordinates within the system
  int local;  // tile index within its system
};

}  // namespace onh
