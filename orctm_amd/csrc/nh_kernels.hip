// nh_kernels.hip -- gfx950 kernels of the nonhydrostatic pressure Poisson solve.
//
// Every launch covers a list of 64x4-column tiles (Tile), possibly from several independent
// systems (Sys); a block finds its system through its tile, so one launch advances a whole batch.
// Reductions are deterministic: per-thread fma accumulation along k, a fixed wavefront butterfly
// (lane 0's value), a fixed order over the 4 wavefronts, block partials summed in tile order by
// the last block of each system (atomic ticket), so a system's result does not depend on what else
// is in the launch.
//
// Floating-point steps that the CPU oracle restates bit for bit are written with explicit fma and
// are compiled with -ffp-contract=off (no other contraction): the assembly of the reference
// operator (mo_petsc.f90:543-782), the FP64 matvec, the inner matvec and the line preconditioner.
#include <math.h>

#include "nh_kernels.h"

namespace onh {

__device__ __forceinline__ double fmad(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of a and b over the block; valid in thread 0.
template <typename T>
__device__ __forceinline__ void block_sum2(T& a, T& b) {
  __shared__ T sh[2 * (NT / 64)];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[w] = a;
    sh[4 + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    b = ((sh[4] + sh[5]) + sh[6]) + sh[7];
  }
}

// Publishes this block's (a, b) (valid in thread 0) and returns true, uniformly, in the last block of
// the system to arrive; that block then holds the system totals in thread 0's a and b.
template <typename T>
__device__ bool last_block2(const Sys& s, int local, T& a, T& b) {
  __shared__ int am_last;
  T* part = reinterpret_cast<T*>(s.partials);
  if (threadIdx.x == 0) {
    part[2 * local] = a;
    part[2 * local + 1] = b;
    __threadfence();
    const unsigned prev = atomicAdd(s.counter, 1u);
    am_last = (prev == (unsigned)(s.ntiles - 1));
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  const volatile T* vp = part;
  T sa = T(0), sb = T(0);
  for (int t = threadIdx.x; t < s.ntiles; t += NT) {
    sa += vp[2 * t];
    sb += vp[2 * t + 1];
  }
  block_sum2(sa, sb);
  if (threadIdx.x == 0) {
    *s.counter = 0u;
    a = sa;
    b = sb;
  }
  return true;
}

// Neighbour map of the DMDA: periodic wrap where cyclic, else validity (outside reads +0).
__device__ __forceinline__ bool map_ij(int& i, int& j, int nx, int ny, int cx, int cy) {
  if (cx) {
    if (i < 0)
      i += nx;
    else if (i >= nx)
      i -= nx;
  }
  if (cy) {
    if (j < 0)
      j += ny;
    else if (j >= ny)
      j -= ny;
  }
  return i >= 0 && i < nx && j >= 0 && j < ny;
}

// ------------------------------------------------------------------------------------------------
// Assembly: compute_matrix (mo_petsc.f90:543-782), one thread per interior column.

#define F2P(a, I, J) (a)[((I)-1) + (size_t)f.IE * ((J)-1)]
#define F2U(a, I, J) (a)[((I)-f.i_start) + (size_t)(f.IE - f.i_start + 1) * ((J)-1)]
#define F2V(a, I, J) (a)[((I)-1) + (size_t)f.IE * ((J)-f.j_start)]
#define F3(a, I, J, K) (a)[((I)-1) + (size_t)f.IE * (((J)-1) + (size_t)f.JE * ((K)-1))]

// dddwo(K) of mo_petsc.f90:557-583 (K = 1..KE+1).
__device__ __forceinline__ double dddwo(const FortranIn& f, int I, int J, int K) {
  const int KE = f.KE;
  if (K == 1) {
    const double d1 = F3(f.ddpo, I, J, 1);
    return d1 == 0.0 ? f.dzw[0] / 2.0 : d1 / 2.0;
  }
  if (K == KE + 1) {
    const double d = F3(f.ddpo, I, J, KE);
    return d == 0.0 ? f.dzw[KE - 1] : d;
  }
  const double dm = F3(f.ddpo, I, J, K - 1), d0 = F3(f.ddpo, I, J, K);
  if (dm == 0.0) return (f.dzw[K - 1] + f.dzw[K - 2]) / 2.0;
  if (d0 == 0.0) return dm;
  return (dm + d0) / 2.0;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_assemble(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                 FortranIn f, int* bad, unsigned long long* smax_bits) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  const bool active = i < s.nx && j < s.ny;
  double smax = 0.0;
  int nbad = 0;
  if (active) {
    const int I = i + 2, J = j + 2, IE1 = f.IE - 1, JE1 = f.JE - 1, KE = f.KE;
    const size_t NP = (size_t)s.plane * s.nz;
    double* A = const_cast<double*>(s.A64);
    double* S = const_cast<double*>(s.S);
    T* C = reinterpret_cast<T*>(const_cast<void*>(s.C));
    uint8_t* wet = const_cast<uint8_t*>(s.wet);
    const double dlxuw = F2U(f.dlxu, I - 1, J), dlxue = F2U(f.dlxu, I, J);
    const double dlyvn = F2V(f.dlyv, I, J - 1), dlyvs = F2V(f.dlyv, I, J);
    const double dlxp = F2P(f.dlxp, I, J), dlyp = F2P(f.dlyp, I, J);
    if (!(dlxuw > 0) || !(dlxue > 0) || !(dlyvn > 0) || !(dlyvs > 0) || !(dlxp > 0) || !(dlyp > 0)) nbad = 1;
    const double xuw = 1.0 / dlxuw, xue = 1.0 / dlxue, xp = 1.0 / dlxp;
    const double yvn = 1.0 / dlyvn, yvs = 1.0 / dlyvs, yp = 1.0 / dlyp;
    double zwK = 1.0 / dddwo(f, I, J, 1);
    for (int K = 1; K <= KE; ++K) {
      const double zwK1 = 1.0 / dddwo(f, I, J, K + 1);
      const double zp = 1.0 / ((1.0 / zwK + 1.0 / zwK1) / 2.0);
      if (!isfinite(zwK) || !isfinite(zwK1) || !isfinite(zp) || !(zp > 0)) nbad = 1;
      double X_W = xuw * xp, X_C = (xuw + xue) * xp, X_E = xue * xp;
      double Y_N = yvn * yp, Y_C = (yvn + yvs) * yp, Y_S = yvs * yp;
      double Z_U = zwK * zp, Z_C = (zwK + zwK1) * zp, Z_D = zwK1 * zp;
      double c[NCOEF] = {0, 0, 0, 0, 0, 0, 0};
      const double w0 = F3(f.weto, I, J, K);
      if (w0 != 0.0 && w0 != 1.0) nbad = 1;
      if (w0 == 0.0) {
        c[0] = -(X_C + Y_C + Z_C);  // land row, mo_petsc.f90:623-626
      } else {
        const double ww = F3(f.weto, I - 1, J, K), we = F3(f.weto, I + 1, J, K);
        const double wn = F3(f.weto, I, J - 1, K), ws = F3(f.weto, I, J + 1, K);
        if (ww == 0.0 && we == 1.0) {  // mo_petsc.f90:630-640
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
        if (wn == 0.0 && ws == 1.0) {  // mo_petsc.f90:643-653
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
        if (K == 1) {  // mo_petsc.f90:657-665
          Z_U = 0.0;
        } else if (K == KE) {
          Z_D = 0.0;
          Z_C = zwK * zp;
        } else if (F3(f.weto, I, J, K + 1) == 0.0) {
          Z_D = 0.0;
          Z_C = zwK * zp;
        }
        if (!f.cyc_x) {  // mo_petsc.f90:669-678
          if (f.open_w && I == 2) X_W = 0.0;
          if (f.open_e && I == IE1) X_E = 0.0;
        }
        if (!f.cyc_y) {  // mo_petsc.f90:680-689
          if (f.open_n && J == 2) Y_N = 0.0;
          if (f.open_s && J == JE1) Y_S = 0.0;
        }
        c[0] = -(X_C + Y_C + Z_C);
        if (ww == 1.0 && (f.cyc_x || I != 2)) c[1] = X_W;  // mo_petsc.f90:692-764
        if (we == 1.0 && (f.cyc_x || I != IE1)) c[2] = X_E;
        if (wn == 1.0 && (f.cyc_y || J != 2)) c[3] = Y_N;
        if (ws == 1.0 && (f.cyc_y || J != JE1)) c[4] = Y_S;
        if (K != 1) c[5] = Z_U;
        if (K != KE && F3(f.weto, I, J, K + 1) == 1.0) c[6] = Z_D;
      }
      const size_t idx = (size_t)(K - 1) * s.plane + (size_t)j * s.pitch + i;
      const double Sv = (dlxp * dlyp) / zp;
#pragma unroll
      for (int q = 0; q < NCOEF; ++q) {
        A[q * NP + idx] = c[q];
        C[q * NP + idx] = (T)(Sv * (-c[q]));
      }
      S[idx] = Sv;
      smax = fmax(smax, Sv);
      wet[idx] = (w0 == 1.0) ? 1 : 0;
      zwK = zwK1;
    }
  }
  // block max and flag (order-independent)
  __shared__ double shm[NT / 64];
  __shared__ int shb[NT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    smax = fmax(smax, __shfl_xor(smax, off, 64));
    nbad |= __shfl_xor(nbad, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    shm[threadIdx.x >> 6] = smax;
    shb[threadIdx.x >> 6] = nbad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = fmax(fmax(shm[0], shm[1]), fmax(shm[2], shm[3]));
    int bflag = shb[0] | shb[1] | shb[2] | shb[3];
    if (bflag) atomicOr(bad, 1);
    if (isfinite(m)) atomicMax(smax_bits, (unsigned long long)__double_as_longlong(m));
  }
}

// Line preconditioner factorisation (per column), restated by orc_line_factor*.
template <typename T>
__global__ __launch_bounds__(NT) void k_line_factor(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const size_t P = s.plane, NP = P * s.nz, col = (size_t)j * s.pitch + i;
  const T* Cc = reinterpret_cast<const T*>(s.C);
  const T* Cu = Cc + 5 * NP;
  const T* Cd = Cc + 6 * NP;
  T* invD = reinterpret_cast<T*>(const_cast<void*>(s.invD));
  T* L = reinterpret_cast<T*>(const_cast<void*>(s.L));
  T Dprev = T(1);
  for (int k = 0; k < s.nz; ++k) {
    const size_t n = col + (size_t)k * P;
    T d = Cc[n], l = T(0);
    if (k > 0) {
      const T e = (Cd[n - P] != T(0) && Cu[n] != T(0)) ? Cu[n] : T(0);
      l = e / Dprev;
      d = fmad(-l, e, d);
    }
    invD[n] = T(1) / d;
    L[n] = l;
    Dprev = d;
  }
}

// ------------------------------------------------------------------------------------------------
// Right-hand side (mo_petsc.f90:376) and initial guess.

__global__ __launch_bounds__(NT) void k_rhs(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                            const double* __restrict__ divg, int IE, int JE) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  double acc = 0.0, dummy = 0.0;
  if (i < s.nx && j < s.ny) {
    const size_t col = (size_t)j * s.pitch + i;
    for (int k = 0; k < s.nz; ++k) {
      const size_t idx = col + (size_t)k * s.plane;
      const double v = divg[(size_t)(i + 1) + (size_t)IE * ((size_t)(j + 1) + (size_t)JE * k)] * (double)s.wet[idx];
      s.b[idx] = v;
      acc = fma(v, v, acc);
    }
  }
  block_sum2(acc, dummy);
  if (last_block2(s, t.local, acc, dummy) && threadIdx.x == 0) *s.nbp = sqrt(acc);
}

__global__ __launch_bounds__(NT) void k_guess(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                              const double* __restrict__ pnh, int IE, int JE) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const size_t col = (size_t)j * s.pitch + i;
  for (int k = 0; k < s.nz; ++k)
    s.p64[col + (size_t)k * s.plane] = pnh[(size_t)(i + 1) + (size_t)IE * ((size_t)(j + 1) + (size_t)JE * k)];
}

// ------------------------------------------------------------------------------------------------
// FP64 reference matvec, restated by orc_matvec64: centre, W, E, N, S, U, D, absent = +0.

__device__ __forceinline__ double apply_a64(const Sys& s, const double* __restrict__ x, size_t NP, int i, int j, int k,
                                            size_t idx) {
  const size_t IP = s.pitch, P = s.plane;
  const int nx = s.nx, ny = s.ny;
  const double xw = i > 0 ? x[idx - 1] : (s.cyc_x ? x[idx + (size_t)(nx - 1)] : 0.0);
  const double xe = i < nx - 1 ? x[idx + 1] : (s.cyc_x ? x[idx - (size_t)(nx - 1)] : 0.0);
  const double xn = j > 0 ? x[idx - IP] : (s.cyc_y ? x[idx + IP * (size_t)(ny - 1)] : 0.0);
  const double xs = j < ny - 1 ? x[idx + IP] : (s.cyc_y ? x[idx - IP * (size_t)(ny - 1)] : 0.0);
  const double xu = k > 0 ? x[idx - P] : 0.0;
  const double xd = k < s.nz - 1 ? x[idx + P] : 0.0;
  const double* A = s.A64;
  double acc = A[idx] * x[idx];
  acc = fma(A[NP + idx], xw, acc);
  acc = fma(A[2 * NP + idx], xe, acc);
  acc = fma(A[3 * NP + idx], xn, acc);
  acc = fma(A[4 * NP + idx], xs, acc);
  acc = fma(A[5 * NP + idx], xu, acc);
  acc = fma(A[6 * NP + idx], xd, acc);
  return acc;
}

__global__ __launch_bounds__(NT) void k_apply64(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                const double* __restrict__ x, double* __restrict__ y) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const size_t NP = (size_t)s.plane * s.nz, col = (size_t)j * s.pitch + i;
  for (int k = 0; k < s.nz; ++k) {
    const size_t idx = col + (size_t)k * s.plane;
    y[idx] = apply_a64(s, x, NP, i, j, k, idx);
  }
}

// ------------------------------------------------------------------------------------------------
// Outer refinement: begin, residual (+ scaled inner right-hand side), refine.

template <typename T>
__global__ void k_begin(const Sys* __restrict__ systems, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const Sys& s = systems[q];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  o->nb = *s.nbp;
  o->smax = *s.smaxp;
  o->rtol = s.rtol;
  o->rtol_inner = s.rtol_inner;
  o->maxit = s.maxit;
  o->max_outer = s.max_outer;
  o->outer = 0;
  o->inner_total = 0;
  o->nr_prev = o->nb;
  o->nr = o->nb;
  o->rel = 0.0;
  o->odone = isfinite(o->nb) ? RUN : NONFINITE;
  in->done = SKIP;
  in->iter = 0;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_residual(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  if (o->odone != RUN) return;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  const double nb = o->nb;
  const bool zero = (nb == 0.0);
  const double gamma = 1.0 / (o->smax * o->nr_prev);
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  double acc = 0.0, dummy = 0.0;
  if (i < s.nx && j < s.ny) {
    const size_t NP = (size_t)s.plane * s.nz, col = (size_t)j * s.pitch + i;
    T* rin = reinterpret_cast<T*>(s.r);
    for (int k = 0; k < s.nz; ++k) {
      const size_t idx = col + (size_t)k * s.plane;
      if (zero) {
        s.p64[idx] = 0.0;
        rin[idx] = T(0);
        continue;
      }
      const double r = s.b[idx] - apply_a64(s, s.p64, NP, i, j, k, idx);
      acc = fma(r, r, acc);
      rin[idx] = (T)((s.S[idx] * (-r)) * gamma);
    }
  }
  block_sum2(acc, dummy);
  if (last_block2(s, t.local, acc, dummy) && threadIdx.x == 0) {
    const double nr = sqrt(acc);
    o->nr = nr;
    int od = RUN;
    if (zero) {
      o->rel = 0.0;
      od = CONVERGED;
    } else {
      const double rel = nr / nb;
      o->rel = rel;
      if (!isfinite(rel))
        od = BREAKDOWN;
      else if (rel <= o->rtol)
        od = CONVERGED;
      else if (o->outer >= o->max_outer || o->inner_total >= o->maxit)
        od = NOTCONV;
      else {
        double tol = 0.5 * o->rtol * nb / nr;
        if (tol < o->rtol_inner) tol = o->rtol_inner;
        in->tol2 = (T)(tol * tol);
        in->maxit = o->maxit - o->inner_total;
        o->inv_gamma = o->smax * o->nr_prev;
        o->nr_prev = nr;
        o->outer += 1;
      }
    }
    o->odone = od;
    if (od != RUN) in->done = SKIP;
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_refine(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const Outer* o = s.outer;
  if (o->odone != RUN) return;
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const double g = o->inv_gamma;
  const T* x = reinterpret_cast<const T*>(s.x);
  const size_t col = (size_t)j * s.pitch + i;
  for (int k = 0; k < s.nz; ++k) {
    const size_t idx = col + (size_t)k * s.plane;
    s.p64[idx] = s.p64[idx] + (double)x[idx] * g;
  }
}

// ------------------------------------------------------------------------------------------------
// Inner PCG. Column kernel (init / update): x, r update, line solve z = M^{-1} r, r.z and r.r.

template <typename T, int NZR, bool INIT>
__global__ __launch_bounds__(NT) void k_column(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (INIT) {
    if (o->odone != RUN) return;
  } else {
    if (in->done != RUN) return;
  }
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  const size_t P = s.plane, col = (size_t)j * s.pitch + i;
  const int nz = s.nz;
  T* __restrict__ x = reinterpret_cast<T*>(s.x);
  T* __restrict__ r = reinterpret_cast<T*>(s.r);
  T* __restrict__ z = reinterpret_cast<T*>(s.z);
  const T* __restrict__ q = reinterpret_cast<const T*>(s.q);
  const T* __restrict__ p = reinterpret_cast<const T*>(in->pbuf ? s.pa : s.pb);
  const T* __restrict__ L = reinterpret_cast<const T*>(s.L);
  const T* __restrict__ invD = reinterpret_cast<const T*>(s.invD);
  const T alpha = INIT ? T(0) : in->alpha;
  T rho = T(0), rr = T(0);
  if (i < s.nx && j < s.ny) {
    T yprev = T(0);
    if constexpr (NZR > 0) {
      T wreg[NZR];
#pragma unroll
      for (int k = 0; k < NZR; ++k) {
        if (k < nz) {
          const size_t idx = col + (size_t)k * P;
          T rv;
          if (INIT) {
            rv = r[idx];
            x[idx] = T(0);
          } else {
            T xv = x[idx];
            rv = r[idx];
            xv = fmad(alpha, p[idx], xv);
            rv = fmad(-alpha, q[idx], rv);
            x[idx] = xv;
            r[idx] = rv;
          }
          rr = fmad(rv, rv, rr);
          const T y = fmad(-L[idx], yprev, rv);
          const T w = y * invD[idx];
          rho = fmad(y, w, rho);
          wreg[k] = w;
          yprev = y;
        }
      }
      T zn = T(0);
#pragma unroll
      for (int k = NZR - 1; k >= 0; --k) {
        if (k < nz) {
          const size_t idx = col + (size_t)k * P;
          T zk;
          if (k == nz - 1)
            zk = wreg[k];
          else
            zk = fmad(-L[idx + P], zn, wreg[k]);
          z[idx] = zk;
          zn = zk;
        }
      }
    } else {
      for (int k = 0; k < nz; ++k) {
        const size_t idx = col + (size_t)k * P;
        T rv;
        if (INIT) {
          rv = r[idx];
          x[idx] = T(0);
        } else {
          T xv = x[idx];
          rv = r[idx];
          xv = fmad(alpha, p[idx], xv);
          rv = fmad(-alpha, q[idx], rv);
          x[idx] = xv;
          r[idx] = rv;
        }
        rr = fmad(rv, rv, rr);
        const T y = fmad(-L[idx], yprev, rv);
        const T w = y * invD[idx];
        rho = fmad(y, w, rho);
        z[idx] = w;
        yprev = y;
      }
      T zn = z[col + (size_t)(nz - 1) * P];
      for (int k = nz - 2; k >= 0; --k) {
        const size_t idx = col + (size_t)k * P;
        const T zk = fmad(-L[idx + P], zn, z[idx]);
        z[idx] = zk;
        zn = zk;
      }
    }
  }
  block_sum2(rho, rr);
  if (last_block2(s, t.local, rho, rr) && threadIdx.x == 0) {
    if (INIT) {
      in->rho = rho;
      in->rr0 = rr;
      in->rr = rr;
      in->iter = 0;
      in->first = 1;
      in->pbuf = 0;
      in->beta = T(0);
      if (rr == T(0))
        in->done = CONVERGED;
      else if (!(rho > T(0)) || !isfinite((double)rho) || !isfinite((double)rr))
        in->done = BREAKDOWN;
      else
        in->done = RUN;
      if (in->done == BREAKDOWN) o->odone = BREAKDOWN;
    } else {
      in->iter += 1;
      o->inner_total += 1;
      in->rr = rr;
      int d = RUN;
      if (rr <= in->tol2 * in->rr0)
        d = CONVERGED;
      else if (!(rho > T(0)) || !isfinite((double)rho) || !isfinite((double)rr))
        d = BREAKDOWN;
      else if (in->iter >= in->maxit)
        d = NOTCONV;
      else {
        in->beta = rho / in->rho;
        in->rho = rho;
      }
      in->first = 0;
      in->pbuf ^= 1;
      in->done = d;
    }
  }
}

// Matvec kernel: p_new = z + beta p_old (or z on the first iteration), q = C p_new, p.q.
// 64x4 tile in LDS, k-marching with p(k-1), p(k), p(k+1) in registers; three LDS planes so one
// barrier per level suffices. Halo: wave 0 the north row, wave 1 the south row, wave 2 the west
// and east columns.
template <typename T>
__global__ __launch_bounds__(NT) void k_spmv(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (in->done != RUN) return;
  const int tx = threadIdx.x & (TX - 1), ty = threadIdx.x >> 6;
  const int i0 = t.bx * TX, j0 = t.by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const bool active = i < nx && j < ny;
  const size_t P = s.plane, IP = s.pitch, NP = P * (size_t)nz;
  const int first = in->first, pb = in->pbuf;
  const T beta = in->beta;
  const T* __restrict__ zv = reinterpret_cast<const T*>(s.z);
  const T* __restrict__ pold = reinterpret_cast<const T*>(pb ? s.pb : s.pa);
  T* __restrict__ pnew = reinterpret_cast<T*>(pb ? s.pa : s.pb);
  T* __restrict__ q = reinterpret_cast<T*>(s.q);
  const T* __restrict__ C = reinterpret_cast<const T*>(s.C);

  __shared__ T lds[3][TY + 2][TX + 2];

  int ci = i, cj = j;
  const bool cval = map_ij(ci, cj, nx, ny, s.cyc_x, s.cyc_y);
  const size_t coff = cval ? (size_t)cj * IP + ci : 0;
  int hx = -1, hy = 0, hi = 0, hj = 0;
  if (ty == 0) {
    hi = i;
    hj = j0 - 1;
    hx = tx + 1;
    hy = 0;
  } else if (ty == 1) {
    hi = i;
    hj = j0 + TY;
    hx = tx + 1;
    hy = TY + 1;
  } else if (ty == 2 && tx < 2 * TY) {
    if (tx < TY) {
      hi = i0 - 1;
      hj = j0 + tx;
      hx = 0;
      hy = tx + 1;
    } else {
      hi = i0 + TX;
      hj = j0 + tx - TY;
      hx = TX + 1;
      hy = tx - TY + 1;
    }
  }
  const bool has_halo = hx >= 0;
  const bool hval = has_halo && map_ij(hi, hj, nx, ny, s.cyc_x, s.cyc_y);
  const size_t hoff = hval ? (size_t)hj * IP + hi : 0;

  auto pv = [&](size_t idx) -> T { return first ? zv[idx] : fmad(beta, pold[idx], zv[idx]); };

  T c0 = cval ? pv(coff) : T(0);
  lds[0][ty + 1][tx + 1] = c0;
  if (has_halo) lds[0][hy][hx] = hval ? pv(hoff) : T(0);
  if (active) pnew[coff] = c0;
  T pm1 = T(0), p0 = c0, dot = T(0);
  __syncthreads();
  for (int k = 0; k < nz; ++k) {
    const size_t ck = (size_t)k * P + coff;
    T cc = T(0), cw = T(0), ce = T(0), cn = T(0), cs = T(0), cu = T(0), cd = T(0);
    if (active) {
      cc = C[ck];
      cw = C[NP + ck];
      ce = C[2 * NP + ck];
      cn = C[3 * NP + ck];
      cs = C[4 * NP + ck];
      cu = C[5 * NP + ck];
      cd = C[6 * NP + ck];
    }
    T pp1 = T(0);
    if (k + 1 < nz) {
      const int b1 = (k + 1) % 3;
      const size_t o1 = (size_t)(k + 1) * P;
      pp1 = cval ? pv(o1 + coff) : T(0);
      lds[b1][ty + 1][tx + 1] = pp1;
      if (has_halo) lds[b1][hy][hx] = hval ? pv(o1 + hoff) : T(0);
      if (active) pnew[o1 + coff] = pp1;
    }
    __syncthreads();
    if (active) {
      const int b0 = k % 3;
      T acc = cc * p0;
      acc = fmad(cw, lds[b0][ty + 1][tx], acc);
      acc = fmad(ce, lds[b0][ty + 1][tx + 2], acc);
      acc = fmad(cn, lds[b0][ty][tx + 1], acc);
      acc = fmad(cs, lds[b0][ty + 2][tx + 1], acc);
      acc = fmad(cu, pm1, acc);
      acc = fmad(cd, pp1, acc);
      q[ck] = acc;
      dot = fmad(p0, acc, dot);
    }
    pm1 = p0;
    p0 = pp1;
  }
  T dummy = T(0);
  block_sum2(dot, dummy);
  if (last_block2(s, t.local, dot, dummy) && threadIdx.x == 0) {
    const T pq = dot;
    const T alpha = in->rho / pq;
    in->pq = pq;
    in->alpha = alpha;
    if (!(pq > T(0)) || !isfinite((double)alpha)) {
      in->done = BREAKDOWN;
      if (in->iter == 0) o->odone = BREAKDOWN;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Launchers.

int pick_nzr(bool fp32, int nz) {
  if (nz <= 32) return 32;
  if (nz <= 64) return 64;
  if (fp32 && nz <= 128) return 128;
  return 0;
}

#define LAUNCH(kern, grid, st, ...)                                \
  do {                                                             \
    if ((grid) <= 0) return hipSuccess;                            \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, st, __VA_ARGS__); \
    return hipGetLastError();                                      \
  } while (0)

hipError_t launch_assemble(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                           const FortranIn& f, int* d_bad, unsigned long long* d_smax_bits) {
  if (fp32) LAUNCH(k_assemble<float>, ntiles, st, d_sys, d_tiles, f, d_bad, d_smax_bits);
  LAUNCH(k_assemble<double>, ntiles, st, d_sys, d_tiles, f, d_bad, d_smax_bits);
}

hipError_t launch_line_factor(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (fp32) LAUNCH(k_line_factor<float>, ntiles, st, d_sys, d_tiles);
  LAUNCH(k_line_factor<double>, ntiles, st, d_sys, d_tiles);
}

hipError_t launch_rhs(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* d_divg, int IE,
                      int JE) {
  LAUNCH(k_rhs, ntiles, st, d_sys, d_tiles, d_divg, IE, JE);
}

hipError_t launch_guess(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* d_pnh, int IE,
                        int JE) {
  LAUNCH(k_guess, ntiles, st, d_sys, d_tiles, d_pnh, IE, JE);
}

hipError_t launch_begin(hipStream_t st, bool fp32, const Sys* d_sys, int nsys) {
  const int g = (nsys + 63) / 64;
  if (fp32)
    hipLaunchKernelGGL(k_begin<float>, dim3(g), dim3(64), 0, st, d_sys, nsys);
  else
    hipLaunchKernelGGL(k_begin<double>, dim3(g), dim3(64), 0, st, d_sys, nsys);
  return hipGetLastError();
}

hipError_t launch_residual(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (fp32) LAUNCH(k_residual<float>, ntiles, st, d_sys, d_tiles);
  LAUNCH(k_residual<double>, ntiles, st, d_sys, d_tiles);
}

hipError_t launch_refine(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (fp32) LAUNCH(k_refine<float>, ntiles, st, d_sys, d_tiles);
  LAUNCH(k_refine<double>, ntiles, st, d_sys, d_tiles);
}

template <bool INIT>
static hipError_t launch_column(hipStream_t st, bool fp32, int nzr, const Sys* d_sys, const Tile* d_tiles,
                                int ntiles) {
  if (fp32) {
    switch (nzr) {
      case 32: LAUNCH((k_column<float, 32, INIT>), ntiles, st, d_sys, d_tiles);
      case 64: LAUNCH((k_column<float, 64, INIT>), ntiles, st, d_sys, d_tiles);
      case 128: LAUNCH((k_column<float, 128, INIT>), ntiles, st, d_sys, d_tiles);
      default: LAUNCH((k_column<float, 0, INIT>), ntiles, st, d_sys, d_tiles);
    }
  }
  switch (nzr) {
    case 32: LAUNCH((k_column<double, 32, INIT>), ntiles, st, d_sys, d_tiles);
    case 64: LAUNCH((k_column<double, 64, INIT>), ntiles, st, d_sys, d_tiles);
    default: LAUNCH((k_column<double, 0, INIT>), ntiles, st, d_sys, d_tiles);
  }
}

hipError_t launch_init(hipStream_t st, bool fp32, int nzr, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  return launch_column<true>(st, fp32, nzr, d_sys, d_tiles, ntiles);
}

hipError_t launch_update(hipStream_t st, bool fp32, int nzr, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  return launch_column<false>(st, fp32, nzr, d_sys, d_tiles, ntiles);
}

hipError_t launch_spmv(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (fp32) LAUNCH(k_spmv<float>, ntiles, st, d_sys, d_tiles);
  LAUNCH(k_spmv<double>, ntiles, st, d_sys, d_tiles);
}

hipError_t launch_apply64(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* x,
                          double* y) {
  LAUNCH(k_apply64, ntiles, st, d_sys, d_tiles, x, y);
}

}  // namespace onh
