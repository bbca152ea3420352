// nh_kernels.hip -- gfx950 kernels of the nonhydrostatic pressure Poisson solve.
//
// Every launch covers a list of 64x4-column tiles (Tile), possibly from several independent
// systems (Sys); a block finds its system through its tile, so one launch advances a whole batch.
// Reductions are deterministic: per-thread fma accumulation along k, a fixed wavefront butterfly
// (lane 0's value), a fixed order over the 4 wavefronts, block partials summed in tile order by a
// one-block-per-system reduction kernel launched after each phase, so a system's result does not
// depend on what else is in the launch.
//
// Floating-point steps that the CPU oracle restates bit for bit are written with explicit fma and
// are compiled with -ffp-contract=off (no other contraction): the assembly of the reference
// operator (mo_petsc.f90:543-782), the FP64 matvec, the inner matvec and the line preconditioner.
#include <math.h>

#include <type_traits>

#include "nh_kernels.h"

namespace onh {

__device__ __forceinline__ double fmad(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return fmaf(a, b, c); }

// Pointers read from a Sys in memory are generic; casting them to the global address space lets the
// compiler use global_load/global_store and count only vector-memory operations in its waits.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> G(const void* p) {
  return (gptr<T>)(const_cast<void*>(p));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of a and b over a block of NW wavefronts; valid in thread 0.
template <typename T, int NW = NT / 64>
__device__ __forceinline__ void block_sum2(T& a, T& b) {
  __shared__ T sh[2 * NW];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[w] = a;
    sh[NW + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T sa = sh[0], sb = sh[NW];
#pragma unroll
    for (int q = 1; q < NW; ++q) {
      sa += sh[q];
      sb += sh[NW + q];
    }
    a = sa;
    b = sb;
  }
}

// Publishes this block's (a, b) (valid in thread 0) into its system's partials slot. The system
// totals are formed by k_reduce, launched right after, in tile order: no atomics, no fences (a
// per-block device-scope fence costs about a third of the bandwidth of these streaming kernels).
template <typename T>
__device__ __forceinline__ void publish2(const Sys& s, int local, T a, T b) {
  if (threadIdx.x == 0) {
    gptr<T> part = G<T>(s.partials);
    part[2 * local] = a;
    part[2 * local + 1] = b;
  }
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
    double m = shm[0];
    int bflag = shb[0];
    for (int q = 1; q < NT / 64; ++q) {
      m = fmax(m, shm[q]);
      bflag |= shb[q];
    }
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
  publish2(s, t.local, acc, dummy);
}

// b <- b * (1 + amplitude * xi), xi ~ N(0,1) per cell from a counter-based SplitMix64 stream
// keyed by (seed, cell index); land cells stay 0. Synthetic time-step forcing for serving loads.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(NT) void k_perturb(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                double amplitude, unsigned long long seed) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  double acc = 0.0, dummy = 0.0;
  if (i < s.nx && j < s.ny) {
    gptr<double> b = G<double>(s.b);
    const size_t col = (size_t)j * s.pitch + i;
    const unsigned long long key = splitmix64(seed);
    for (int k = 0; k < s.nz; ++k) {
      const size_t idx = col + (size_t)k * s.plane;
      const unsigned long long n = ((unsigned long long)k * s.ny + j) * s.nx + i;
      const unsigned long long a = splitmix64(key + 2 * n), c = splitmix64(key + 2 * n + 1);
      const double u1 = 1.0 - (double)(a >> 11) * (1.0 / 9007199254740992.0);
      const double u2 = (double)(c >> 11) * (1.0 / 9007199254740992.0);
      const double xi = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      const double v = b[idx] * (1.0 + amplitude * xi);
      b[idx] = v;
      acc = fma(v, v, acc);
    }
  }
  block_sum2(acc, dummy);
  publish2(s, t.local, acc, dummy);
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

__device__ __forceinline__ double apply_a64(const Sys& s, gptr<const double> x, size_t NP, int i, int j, int k,
                                            size_t idx) {
  const size_t IP = s.pitch, P = s.plane;
  const int nx = s.nx, ny = s.ny;
  // neighbour indices with periodic wrap; absent neighbours read index idx and are replaced by +0
  const bool hw = i > 0 || s.cyc_x, he = i < nx - 1 || s.cyc_x;
  const bool hn = j > 0 || s.cyc_y, hs = j < ny - 1 || s.cyc_y;
  const size_t iw = i > 0 ? idx - 1 : (s.cyc_x ? idx + (size_t)(nx - 1) : idx);
  const size_t ie = i < nx - 1 ? idx + 1 : (s.cyc_x ? idx - (size_t)(nx - 1) : idx);
  const size_t in_ = j > 0 ? idx - IP : (s.cyc_y ? idx + IP * (size_t)(ny - 1) : idx);
  const size_t is = j < ny - 1 ? idx + IP : (s.cyc_y ? idx - IP * (size_t)(ny - 1) : idx);
  const size_t iu = k > 0 ? idx - P : idx;
  const size_t id = k < s.nz - 1 ? idx + P : idx;
  const double x0 = x[idx];
  const double vw = x[iw], ve = x[ie], vn = x[in_], vs = x[is], vu = x[iu], vd = x[id];
  const double xw = hw ? vw : 0.0, xe = he ? ve : 0.0, xn = hn ? vn : 0.0, xs = hs ? vs : 0.0;
  const double xu = k > 0 ? vu : 0.0, xd = k < s.nz - 1 ? vd : 0.0;
  gptr<const double> A = G<const double>(s.A64);
  double acc = A[idx] * x0;
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
    y[idx] = apply_a64(s, G<const double>(x), NP, i, j, k, idx);
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
    gptr<T> rin = G<T>(s.r);
#pragma unroll 2
    for (int k = 0; k < s.nz; ++k) {
      const size_t idx = col + (size_t)k * s.plane;
      if (zero) {
        s.p64[idx] = 0.0;
        rin[idx] = T(0);
        continue;
      }
      const double r = G<const double>(s.b)[idx] - apply_a64(s, G<const double>(s.p64), NP, i, j, k, idx);
      acc = fma(r, r, acc);
      rin[idx] = (T)((G<const double>(s.S)[idx] * (-r)) * gamma);
    }
  }
  block_sum2(acc, dummy);
  publish2(s, t.local, acc, dummy);
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
  gptr<const T> x = G<const T>(s.x);
  gptr<double> p64 = G<double>(s.p64);
  const size_t col = (size_t)j * s.pitch + i;
#pragma unroll 4
  for (int k = 0; k < s.nz; ++k) {
    const size_t idx = col + (size_t)k * s.plane;
    p64[idx] = p64[idx] + (double)x[idx] * g;
  }
}

// ------------------------------------------------------------------------------------------------
// Inner PCG, two kernels per iteration.
//
//   k_fwd  (per column, top-down):   x += alpha p; r -= alpha q; forward elimination of the line
//          preconditioner  y_k = r_k - l_k y_{k-1},  w_k = y_k / D_k;  r.z = sum y_k w_k,  r.r.
//          Reads x p r q l invD, writes x r w: 9 words per cell.
//   k_bwd  (64x4 tile, bottom-up):   back substitution z_k = w_k - l_{k+1} z_{k+1} fused with
//          p = z + beta p_old, q = C p (7-point, horizontal neighbours through LDS) and p.q.
//          Reads w l p_old + 7 coefficients, writes p q: 12 words per cell.
// The halo lanes of k_bwd run the back-substitution recurrence of their halo columns themselves, so
// z never goes to memory and no grid-wide step separates the two halves of the line solve.

template <typename T, bool INIT>
__global__ __launch_bounds__(NT) void k_fwd(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
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
  gptr<T> x = G<T>(s.x);
  gptr<T> r = G<T>(s.r);
  gptr<T> w = G<T>(s.z);
  gptr<const T> q = G<const T>(s.q);
  gptr<const T> p = G<const T>(in->pbuf ? s.pa : s.pb);
  gptr<const T> L = G<const T>(s.L);
  gptr<const T> invD = G<const T>(s.invD);
  const T alpha = INIT ? T(0) : in->alpha;
  T rho = T(0), rr = T(0);
  if (i < s.nx && j < s.ny) {
    // Batches of U levels in two register sets (ping-pong): the loads of batch b+1 are issued before
    // batch b is computed and stored, so stores (which may alias later loads as far as the compiler
    // knows) never sit between a load and its use, and no prefetched value has to be moved.
    constexpr int U = 4;
    struct Batch {
      T x[U], p[U], r[U], q[U], l[U], d[U];
    };
    Batch b0, b1;
    auto load = [&](int k0, Batch& b) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u < nz ? k0 + u : nz - 1;
        const size_t idx = col + (size_t)k * P;
        b.r[u] = r[idx];
        b.l[u] = L[idx];
        b.d[u] = invD[idx];
        if (!INIT) {
          b.x[u] = x[idx];
          b.p[u] = p[idx];
          b.q[u] = q[idx];
        }
      }
    };
    T yprev = T(0);
    auto compute = [&](int k0, const Batch& b) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k0 + u < nz) {
          const size_t idx = col + (size_t)(k0 + u) * P;
          T rk = b.r[u];
          if (INIT) {
            x[idx] = T(0);
          } else {
            x[idx] = fmad(alpha, b.p[u], b.x[u]);
            rk = fmad(-alpha, b.q[u], rk);
            r[idx] = rk;
          }
          rr = fmad(rk, rk, rr);
          const T y = fmad(-b.l[u], yprev, rk);
          const T wv = y * b.d[u];
          rho = fmad(y, wv, rho);
          w[idx] = wv;
          yprev = y;
        }
      }
    };
    load(0, b0);
    for (int k0 = 0; k0 < nz; k0 += 2 * U) {
      load(k0 + U, b1);
      compute(k0, b0);
      if (k0 + U >= nz) break;
      load(k0 + 2 * U, b0);
      compute(k0 + U, b1);
    }
  }
  block_sum2(rho, rr);
  publish2(s, t.local, rho, rr);
}

// Threads of the matvec kernel: TY wavefronts own the tile rows (one column each, centre slots);
// three more wavefronts own the halo: the north row, the south row and the 2*TY west/east cells.
// Every thread runs the back-substitution recurrence of exactly one column, so no wavefront does
// both centre and halo work between two barriers.
constexpr int NW_BWD = TY + 3;
constexpr int NT_BWD = 64 * NW_BWD;

// raw != 0 (verification only): w is taken as z itself (l treated as 0), so q = C w.
template <typename T>
__global__ __launch_bounds__(NT_BWD) void k_bwd(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                int raw) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (in->done != RUN) return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i0 = t.bx * TX, j0 = t.by * TY;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const size_t P = s.plane, IP = s.pitch, NP = P * (size_t)nz;
  const bool first = in->first != 0;
  const int pb = in->pbuf;
  const T beta = in->beta;
  gptr<const T> w = G<const T>(s.z);
  gptr<const T> L = G<const T>(s.L);
  gptr<const T> pold = G<const T>(pb ? s.pb : s.pa);
  gptr<T> pnew = G<T>(pb ? s.pa : s.pb);
  gptr<T> q = G<T>(s.q);
  gptr<const T> C = G<const T>(s.C);

  __shared__ T lds[3][TY + 2][TX + 2];

  // this thread's slot in the LDS plane and the column it follows
  int gi, gj, sx, sy;
  bool slot = true;
  const bool centre = ty < TY;
  if (centre) {
    gi = i0 + tx;
    gj = j0 + ty;
    sx = tx + 1;
    sy = ty + 1;
  } else if (ty == TY) {
    gi = i0 + tx;
    gj = j0 - 1;
    sx = tx + 1;
    sy = 0;
  } else if (ty == TY + 1) {
    gi = i0 + tx;
    gj = j0 + TY;
    sx = tx + 1;
    sy = TY + 1;
  } else {
    slot = tx < 2 * TY;
    const int e = tx < TY ? 0 : 1, r = tx < TY ? tx : tx - TY;
    gi = e ? i0 + TX : i0 - 1;
    gj = j0 + (slot ? r : 0);
    sx = e ? TX + 1 : 0;
    sy = (slot ? r : 0) + 1;
  }
  const bool active = centre && gi < nx && gj < ny;
  int mi = gi, mj = gj;
  const bool val = slot && map_ij(mi, mj, nx, ny, s.cyc_x, s.cyc_y);
  const size_t off = val ? (size_t)mj * IP + mi : 0;

  // Software pipeline over levels (bottom-up) in two register sets used alternately: while level k
  // is finished with set `cur` (p at level k-1 from its w/l/p_old, q at level k from its
  // coefficients), set `nxt` receives w/l/p_old of level k-2 and the coefficients of level k-1.
  struct Stage {
    T w, l, p;
    T c[NCOEF];
  };
  T zc = T(0), lc = T(0);  // z and l of this thread's column at the level below
  auto load_a = [&](int m, Stage& st) {
    const size_t om = (size_t)(m > 0 ? m : 0) * P + off;
    st.w = w[om];
    st.l = L[om];
    st.p = pold[om];
  };
  auto load_c = [&](int m, Stage& st) {
    if (!centre) return;
    const size_t cm = (size_t)(m > 0 ? m : 0) * P + off;
#pragma unroll
    for (int c = 0; c < NCOEF; ++c) st.c[c] = C[c * NP + cm];
  };
  // back substitution and search-direction update of level m for this thread's column
  auto level = [&](int m, const Stage& st) -> T {
    const T zm = fmad(-lc, zc, st.w);
    const T pm = first ? zm : fmad(beta, st.p, zm);
    zc = zm;
    lc = raw ? T(0) : st.l;
    const T pout = val ? pm : T(0);
    if (slot) lds[m % 3][sy][sx] = pout;
    if (active) pnew[(size_t)m * P + off] = pm;
    return pout;
  };
  T p0 = T(0), pp1 = T(0), dot = T(0);
  auto step = [&](int k, Stage& cur, Stage& nxt) {
    load_a(k - 2, nxt);
    load_c(k - 1, nxt);
    const T pm1 = k > 0 ? level(k - 1, cur) : T(0);
    __syncthreads();
    if (centre) {
      const int b0 = k % 3;
      T acc = cur.c[0] * p0;
      acc = fmad(cur.c[1], lds[b0][sy][sx - 1], acc);
      acc = fmad(cur.c[2], lds[b0][sy][sx + 1], acc);
      acc = fmad(cur.c[3], lds[b0][sy - 1][sx], acc);
      acc = fmad(cur.c[4], lds[b0][sy + 1][sx], acc);
      acc = fmad(cur.c[5], pm1, acc);
      acc = fmad(cur.c[6], pp1, acc);
      if (active) {
        q[(size_t)k * P + off] = acc;
        dot = fmad(p0, acc, dot);
      }
    }
    pp1 = p0;
    p0 = pm1;
  };
  Stage s0, s1;
  load_a(nz - 1, s1);
  load_a(nz - 2, s0);
  load_c(nz - 1, s0);
  p0 = level(nz - 1, s1);
  __syncthreads();
  for (int k = nz - 1; k >= 0; k -= 2) {
    step(k, s0, s1);
    if (k == 0) break;
    step(k - 1, s1, s0);
  }
  T dummy = T(0);
  block_sum2<T, NW_BWD>(dot, dummy);
  publish2(s, t.local, dot, dummy);
}

// ------------------------------------------------------------------------------------------------
// System totals and scalar recurrences: one block per system, partials summed in tile order.

enum RedMode { RED_NB = 0, RED_RES = 1, RED_INIT = 2, RED_FWD = 3, RED_BWD = 4 };

template <typename T, int MODE>
__global__ __launch_bounds__(NT) void k_reduce(const Sys* __restrict__ systems) {
  const Sys& s = systems[blockIdx.x];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if ((MODE == RED_RES || MODE == RED_INIT) && o->odone != RUN) return;
  if ((MODE == RED_FWD || MODE == RED_BWD) && in->done != RUN) return;
  // the rhs and residual kernels publish FP64 partials; the inner kernels publish T partials
  using Acc = typename std::conditional<(MODE == RED_NB || MODE == RED_RES), double, T>::type;
  gptr<const Acc> part = G<const Acc>(s.partials);
  Acc a = Acc(0), b = Acc(0);
  for (int t = threadIdx.x; t < s.ntiles; t += NT) {
    a += part[2 * t];
    b += part[2 * t + 1];
  }
  block_sum2(a, b);
  if (threadIdx.x != 0) return;
  if (MODE == RED_NB) {
    *s.nbp = sqrt((double)a);
  } else if (MODE == RED_RES) {
    // FP64 residual norm: convergence test and set-up of the next inner solve
    const double nb = o->nb;
    const bool zero = (nb == 0.0);
    const double nr = sqrt((double)a);
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
  } else if (MODE == RED_INIT) {
    const T rho = a, rr = b;
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
  } else if (MODE == RED_FWD) {
    const T rho = a, rr = b;
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
  } else {  // RED_BWD
    const T pq = a;
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
// Plain copy (16 B per lane, grid-stride): the device's achievable streaming bandwidth, and the
// known-size reference for converting HBM byte counters of the solver kernels.

__global__ __launch_bounds__(256) void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n4; i += st) b[i] = a[i];
}

// ------------------------------------------------------------------------------------------------
// Launchers.

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

template <typename T, int MODE>
static hipError_t launch_reduce(hipStream_t st, const Sys* d_sys, int nsys) {
  hipLaunchKernelGGL((k_reduce<T, MODE>), dim3(nsys), dim3(NT), 0, st, d_sys);
  return hipGetLastError();
}

#define PHASE(kern, red, grid, st, ...)                                        \
  do {                                                                         \
    if ((grid) <= 0) return hipSuccess;                                        \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, st, __VA_ARGS__);         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) return e_;                                           \
    return red;                                                                \
  } while (0)

hipError_t launch_rhs(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                      const double* d_divg, int IE, int JE) {
  PHASE(k_rhs, (launch_reduce<double, RED_NB>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles, d_divg, IE, JE);
}

hipError_t launch_perturb(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                          double amplitude, unsigned long long seed) {
  PHASE(k_perturb, (launch_reduce<double, RED_NB>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles, amplitude, seed);
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

hipError_t launch_residual(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys) {
  if (fp32) PHASE(k_residual<float>, (launch_reduce<float, RED_RES>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles);
  PHASE(k_residual<double>, (launch_reduce<double, RED_RES>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles);
}

hipError_t launch_refine(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (fp32) LAUNCH(k_refine<float>, ntiles, st, d_sys, d_tiles);
  LAUNCH(k_refine<double>, ntiles, st, d_sys, d_tiles);
}

hipError_t launch_init(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys) {
  if (fp32)
    PHASE((k_fwd<float, true>), (launch_reduce<float, RED_INIT>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles);
  PHASE((k_fwd<double, true>), (launch_reduce<double, RED_INIT>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles);
}

hipError_t launch_update(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys) {
  if (fp32)
    PHASE((k_fwd<float, false>), (launch_reduce<float, RED_FWD>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles);
  PHASE((k_fwd<double, false>), (launch_reduce<double, RED_FWD>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles);
}

hipError_t launch_spmv(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys,
                       int raw) {
  if (ntiles <= 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL(k_bwd<float>, dim3(ntiles), dim3(NT_BWD), 0, st, d_sys, d_tiles, raw);
  else
    hipLaunchKernelGGL(k_bwd<double>, dim3(ntiles), dim3(NT_BWD), 0, st, d_sys, d_tiles, raw);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return fp32 ? launch_reduce<float, RED_BWD>(st, d_sys, nsys) : launch_reduce<double, RED_BWD>(st, d_sys, nsys);
}

hipError_t launch_copy4(hipStream_t st, const void* src, void* dst, size_t bytes) {
  const size_t n4 = bytes / 16;
  hipLaunchKernelGGL(k_copy4, dim3(4096), dim3(256), 0, st, (const float4*)src, (float4*)dst, n4);
  return hipGetLastError();
}

hipError_t launch_apply64(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* x,
                          double* y) {
  LAUNCH(k_apply64, ntiles, st, d_sys, d_tiles, x, y);
}

}  // namespace onh
