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
#include "nh_flags.h"

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

// Storage of the FP32 line-preconditioner factors: IEEE fp16 bits, normal range only (the rules of
// the oracle's factor_round32): l with |l| < 2^-14 is stored as 0; invD is scaled by a power of two
// per column (largest scaled value in [2^14, 2^15)) with magnitudes raised to at least 2^-14. The
// rounding is done with integer operations (ties to even) so that it is the oracle's to the bit;
// decoding is the exact fp16 -> fp32 conversion.
__device__ __forceinline__ uint16_t half_bits(float v) {
  const uint32_t u0 = __float_as_uint(v);
  const uint32_t sgn = (u0 >> 16) & 0x8000u;
  const float a = fabsf(v);
  if (!(a < 65520.0f)) return isnan(v) ? (uint16_t)0x7E00u : (uint16_t)(sgn | 0x7C00u);
  if (a == 0.0f) return (uint16_t)sgn;
  const uint32_t u = (u0 + 0xFFFu + ((u0 >> 13) & 1u)) & 0xFFFFE000u;
  const uint32_t e = ((u >> 23) & 0xFFu) - 112u;  // fp32 exponent bias 127 -> fp16 bias 15
  return (uint16_t)(sgn | (e << 10) | ((u >> 13) & 0x3FFu));
}

__device__ __forceinline__ float half_val(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

constexpr float kHalfMin = 6.103515625e-05f;  // 2^-14

// Streaming accesses of the inner kernels: every value is touched once per launch and re-read only
// by the next launch, a gigabyte or more later, so they carry the non-temporal hint (no cache
// residency is expected; the stores do not evict lines other blocks still read).
template <typename T>
__device__ __forceinline__ T ld_nt(gptr<const T> p, size_t i) {
  return __builtin_nontemporal_load(p + i);
}
template <typename T>
__device__ __forceinline__ void st_nt(gptr<T> p, size_t i, T v) {
  __builtin_nontemporal_store(v, p + i);
}

// Buffer-descriptor access (the matvec kernel's column march). A descriptor covers one array; a lane
// addresses it by a 32-bit byte offset (its column's offset plus the level's). A lane that must not
// touch memory takes an offset at or past kOob, beyond every array of at most kBufMaxBytes: the
// hardware's range check then returns 0 for its load and drops its store. A whole wavefront is masked
// by a descriptor of size 0. So every wavefront issues the same memory instructions on every path, and
// the compiler's wait counts keep the loads of the next level in flight instead of draining them where
// branches around loads and stores join.
constexpr uint32_t kOob = 0x80000000u;
static_assert(kBufMaxBytes <= kOob, "offsets of valid lanes stay below kOob");
constexpr int kBufWord3 = 0x00020000;  // gfx9 descriptor word 3: 32-bit data format, untyped access
using Rsrc = __amdgpu_buffer_rsrc_t;
enum : int { kAuxNone = 0, kAuxNt = 2 };  // cache policy bits: none, non-temporal

__device__ __forceinline__ Rsrc buf_of(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufWord3);
}
template <typename E, int AUX>
__device__ __forceinline__ E bld(Rsrc r, uint32_t vo) {
  if constexpr (sizeof(E) == 2) {
    return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b16(r, vo, 0, AUX));
  } else if constexpr (sizeof(E) == 4) {
    return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, vo, 0, AUX));
  } else {
    static_assert(sizeof(E) == 8, "2-, 4- or 8-byte elements");
    return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, vo, 0, AUX));
  }
}
template <typename E>
__device__ __forceinline__ void bst_nt(Rsrc r, uint32_t vo, E v) {
  if constexpr (sizeof(E) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, 0, kAuxNt);
  } else {
    static_assert(sizeof(E) == 8, "4- or 8-byte elements");
    typedef unsigned int U2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, vo, 0, kAuxNt);
  }
}

// Factor storage per inner precision: fp16 bits for FP32, the value itself for FP64.
template <typename T>
struct Fac {
  using S = T;
  __device__ static T dec(T v) { return v; }
  __device__ static T ld(gptr<const S> p, size_t i) { return p[i]; }
  __device__ static T ld_nt(gptr<const S> p, size_t i) { return onh::ld_nt(p, i); }
};
template <>
struct Fac<float> {
  using S = uint16_t;
  __device__ static float dec(uint16_t v) { return half_val(v); }
  __device__ static float ld(gptr<const uint16_t> p, size_t i) { return half_val(p[i]); }
  __device__ static float ld_nt(gptr<const uint16_t> p, size_t i) { return half_val(onh::ld_nt(p, i)); }
};

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

// Publishes the per-row sums of (a, b) of a block whose wavefront w (w < H) holds row j0 + w of its
// tile, as totals over fixed groups of RG = 4 consecutive rows: group g = j / 4 of x-tile bx goes to
// slot g * tiles_x + bx of the system's partials, as ((r0 + r1) + (r2 + r3)) of its row sums. The
// system totals, formed by k_reduce over the slots in a fixed order, are therefore the same bits for
// every tile height H (any multiple of 4), so the tile shape can be tuned without changing results.
// No atomics and no fences (a per-block device-scope fence costs about a third of the bandwidth of
// these streaming kernels). Wavefronts w >= H (halo waves) contribute nothing.
template <typename T, int H, int NW>
__device__ __forceinline__ void publish_rows(const Sys& s, int bx, int by, T a, T b) {
  static_assert(H % RG == 0 && H <= NW, "row groups must not straddle tiles");
  __shared__ T sh[2 * H];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0 && w < H) {
    sh[w] = a;
    sh[H + w] = b;
  }
  __syncthreads();
  if (threadIdx.x < H / RG) {
    const int g = threadIdx.x, grow = by * (H / RG) + g;
    if (grow * RG < s.ny) {
      const T sa = (sh[4 * g] + sh[4 * g + 1]) + (sh[4 * g + 2] + sh[4 * g + 3]);
      const T sb = (sh[H + 4 * g] + sh[H + 4 * g + 1]) + (sh[H + 4 * g + 2] + sh[H + 4 * g + 3]);
      gptr<T> part = G<T>(s.partials);
      const size_t slot = (size_t)grow * s.tiles_x + bx;
      part[2 * slot] = sa;
      part[2 * slot + 1] = sb;
    }
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

// One row of compute_matrix rebuilt from the compact form, in T arithmetic (double: the operations
// of the assembly, bit for bit; float: the inner operator). c7: centre, W, E, N, S, U, D.
// flags: bit 0 wet; bits 1-6 coupling present W E N S U D; bits 7-8 X_C variant (full, XUe*XP,
// XUw*XP, 0); bits 9-10 Y_C variant (full, YVs*YP, YVn*YP, 0); bit 11 Z_C variant (full, ZW*ZP).
template <typename T>
__device__ __forceinline__ void row_coef(const T* cc, T zwk, T zwk1, T zpk, unsigned f, T* c7) {
  const T xw = cc[0], xe = cc[1], yn = cc[2], ys = cc[3];
  const unsigned xc = (f >> 7) & 3u, yc = (f >> 9) & 3u;
  const T X_C = xc == 0 ? cc[4] : xc == 1 ? xe : xc == 2 ? xw : T(0);
  const T Y_C = yc == 0 ? cc[5] : yc == 1 ? ys : yc == 2 ? yn : T(0);
  const T Z_C = ((f >> 11) & 1u) ? zwk * zpk : (zwk + zwk1) * zpk;
  c7[0] = -(X_C + Y_C + Z_C);
  c7[1] = (f & (1u << 1)) ? xw : T(0);
  c7[2] = (f & (1u << 2)) ? xe : T(0);
  c7[3] = (f & (1u << 3)) ? yn : T(0);
  c7[4] = (f & (1u << 4)) ? ys : T(0);
  c7[5] = (f & (1u << 5)) ? zwk * zpk : T(0);
  c7[6] = (f & (1u << 6)) ? zwk1 * zpk : T(0);
}

// ZW/ZP of one column from the per-level table and the column's 3-level window, and the column's flag
// bits from its wet depths (Sys::tab).
template <typename T>
struct ZCol {
  gptr<const T> tw, tp;
  int kx;
  T w0, w1, w2, p0, p1, p2;
  int kc, kw, ke, kn, ks;
  unsigned edge;
  __device__ void load(gptr<const T> tw_, gptr<const T> tp_, const int* xk, gptr<const T> xw, gptr<const T> xp,
                       const uint32_t* xf, size_t P, size_t off) {
    tw = tw_;
    tp = tp_;
    kx = xk[off];
    w0 = xw[off];
    w1 = xw[P + off];
    w2 = xw[2 * P + off];
    p0 = xp[off];
    p1 = xp[P + off];
    p2 = xp[2 * P + off];
    const uint32_t a = xf[off], b = xf[P + off], c = xf[2 * P + off];
    kc = (int)(a & 0xffffu);
    kw = (int)(a >> 16);
    ke = (int)(b & 0xffffu);
    kn = (int)(b >> 16);
    ks = (int)(c & 0xffffu);
    edge = c >> 16;
  }
  // flag bits of level m (0-based) of a column of nz levels
  __device__ unsigned flag(int m, int nz) const { return derive_flags(m + 1, nz, kc, kw, ke, kn, ks, edge); }
  __device__ T zw(int m) const {
    const int d = m - kx;
    return d == 0 ? w0 : d == 1 ? w1 : d == 2 ? w2 : tw[m];
  }
  __device__ T zp(int m) const {
    const int d = m - kx;
    return d == 0 ? p0 : d == 1 ? p1 : d == 2 ? p2 : tp[m];
  }
  // the window's value at level m where the window covers m, else t (the table's entry)
  __device__ T zw_or(int m, T t) const {
    const int d = m - kx;
    return d == 0 ? w0 : d == 1 ? w1 : d == 2 ? w2 : t;
  }
  __device__ T zp_or(int m, T t) const {
    const int d = m - kx;
    return d == 0 ? p0 : d == 1 ? p1 : d == 2 ? p2 : t;
  }
};

// The wet column's ZW/ZP (mo_petsc.f90:557-590 with DDPO = DZW on every level), with the assembly's
// expressions, so that full columns reproduce the assembled values bit for bit. One thread.
template <typename T>
__global__ void k_ztab(const Sys* __restrict__ systems, FortranIn f) {
  const Sys& s = systems[0];
  if (threadIdx.x != 0) return;
  gptr<double> tw = G<double>(s.tw);
  gptr<double> tp = G<double>(s.tp);
  gptr<T> twi = G<T>(s.tw_in);
  gptr<T> tpi = G<T>(s.tp_in);
  const bool copy_in = (const void*)s.tw_in != (const void*)s.tw;
  const int KE = f.KE;
  double zwK = 1.0 / (f.dzw[0] / 2.0);
  tw[0] = zwK;
  if (copy_in) twi[0] = (T)zwK;
  for (int K = 1; K <= KE; ++K) {
    const double dd = K < KE ? (f.dzw[K - 1] + f.dzw[K]) / 2.0 : f.dzw[KE - 1];
    const double zwK1 = 1.0 / dd;
    const double zpK = 1.0 / ((1.0 / zwK + 1.0 / zwK1) / 2.0);
    tw[K] = zwK1;
    tp[K - 1] = zpK;
    if (copy_in) {
      twi[K] = (T)zwK1;
      tpi[K - 1] = (T)zpK;
    }
    zwK = zwK1;
  }
}

// Per column: the levels where the assembled ZW/ZP (either precision) differ from the table, which
// must fit in a window of 3 consecutive levels; the window's values. *bad set if some column does not fit.
template <typename T>
__global__ __launch_bounds__(NT) void k_zwindow(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                int* bad) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const int nz = s.nz;
  const size_t P = s.plane, col = (size_t)j * s.pitch + i;
  gptr<const double> zw = G<const double>(s.zw), zp = G<const double>(s.zp), tw = G<const double>(s.tw),
                     tp = G<const double>(s.tp);
  gptr<const T> zwi = G<const T>(s.zw_in), zpi = G<const T>(s.zp_in), twi = G<const T>(s.tw_in),
                tpi = G<const T>(s.tp_in);
  auto same64 = [](double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); };
  auto sameT = [](T a, T b) {
    if (sizeof(T) == 8) return __double_as_longlong((double)a) == __double_as_longlong((double)b);
    return __float_as_uint((float)a) == __float_as_uint((float)b);
  };
  int lo = nz + 1, hi = -1;
  for (int m = 0; m <= nz; ++m) {
    const size_t o = (size_t)m * P + col;
    bool d = !same64(zw[o], tw[m]) || !sameT(zwi[o], twi[m]);
    if (m < nz) d = d || !same64(zp[o], tp[m]) || !sameT(zpi[o], tpi[m]);
    if (d) {
      lo = lo < m ? lo : m;
      hi = m;
    }
  }
  int kx = 0;
  if (hi >= 0) {
    kx = lo;
    const int top = nz >= 2 ? nz - 2 : 0;  // keep the window within ZW's nz+1 levels
    if (kx > top) kx = top;
    if (hi > kx + 2) atomicOr(bad, 1);
  }
  int* xk = const_cast<int*>(s.xk);
  gptr<double> xw = G<double>(s.xw), xp = G<double>(s.xp);
  gptr<T> xwi = G<T>(s.xw_in), xpi = G<T>(s.xp_in);
  const bool copy_in = (const void*)s.xw_in != (const void*)s.xw;
  xk[col] = kx;
  for (int c = 0; c < 3; ++c) {
    const int m = kx + c;
    const double w = m <= nz ? zw[(size_t)m * P + col] : 0.0;
    const double p = m < nz ? zp[(size_t)m * P + col] : 0.0;
    xw[c * P + col] = w;
    xp[c * P + col] = p;
    if (copy_in) {
      xwi[c * P + col] = m <= nz ? zwi[(size_t)m * P + col] : T(0);
      xpi[c * P + col] = m < nz ? zpi[(size_t)m * P + col] : T(0);
    }
  }
}

// Per column: the wet depths of the column and of its 4 neighbours from WETO (halo included) and the
// edge bits, packed into Sys::xf (nh_flags.h); every level's derived flag bits compared with the
// assembled flag plane. *bad set if any cell differs (a wet cell under a dry one, a WETO value other
// than 0 or 1).
__global__ __launch_bounds__(NT) void k_zflags(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                               FortranIn f, int* bad) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  uint32_t w[3];
  column_flag_words(f.weto, f.IE, f.JE, f.KE, i + 2, j + 2, f.cyc_x, f.cyc_y, w);
  const size_t P = s.plane, col = (size_t)j * s.pitch + i;
  uint32_t* xf = const_cast<uint32_t*>(s.xf);
  xf[col] = w[0];
  xf[P + col] = w[1];
  xf[2 * P + col] = w[2];
  gptr<const uint16_t> flg = G<const uint16_t>(s.flg);
  bool diff = false;
  for (int K = 1; K <= f.KE; ++K) diff = diff || flg[(size_t)(K - 1) * P + col] != flags_from_words(K, f.KE, w);
  if (diff) atomicOr(bad, 1);
}

// Assembly into the compact form. ZW/ZP follow mo_petsc.f90:557-590, the per-column products
// mo_petsc.f90:593-599,610-615, the flag bits the branches of mo_petsc.f90:623-765.
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
    const size_t P = s.plane, col = (size_t)j * s.pitch + i;
    gptr<double> zw = G<double>(s.zw);
    gptr<double> zp = G<double>(s.zp);
    gptr<uint16_t> flg = G<uint16_t>(s.flg);
    gptr<double> cp = G<double>(s.col);
    gptr<T> zwi = G<T>(s.zw_in);
    gptr<T> zpi = G<T>(s.zp_in);
    gptr<T> cpi = G<T>(s.col_in);
    const bool copy_in = (const void*)s.zw_in != (const void*)s.zw;
    const double dlxuw = F2U(f.dlxu, I - 1, J), dlxue = F2U(f.dlxu, I, J);
    const double dlyvn = F2V(f.dlyv, I, J - 1), dlyvs = F2V(f.dlyv, I, J);
    const double dlxp = F2P(f.dlxp, I, J), dlyp = F2P(f.dlyp, I, J);
    if (!(dlxuw > 0) || !(dlxue > 0) || !(dlyvn > 0) || !(dlyvs > 0) || !(dlxp > 0) || !(dlyp > 0)) nbad = 1;
    const double xuw = 1.0 / dlxuw, xue = 1.0 / dlxue, xp = 1.0 / dlxp;
    const double yvn = 1.0 / dlyvn, yvs = 1.0 / dlyvs, yp = 1.0 / dlyp;
    const double cc[7] = {xuw * xp, xue * xp, yvn * yp, yvs * yp, (xuw + xue) * xp, (yvn + yvs) * yp, dlxp * dlyp};
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      cp[q * P + col] = cc[q];
      if (copy_in) cpi[q * P + col] = (T)cc[q];
    }
    double zwK = 1.0 / dddwo(f, I, J, 1);
    zw[col] = zwK;
    if (copy_in) zwi[col] = (T)zwK;
    for (int K = 1; K <= KE; ++K) {
      const size_t idx = (size_t)(K - 1) * P + col;
      const double zwK1 = 1.0 / dddwo(f, I, J, K + 1);
      const double zpK = 1.0 / ((1.0 / zwK + 1.0 / zwK1) / 2.0);
      if (!isfinite(zwK) || !isfinite(zwK1) || !isfinite(zpK) || !(zpK > 0)) nbad = 1;
      zw[idx + P] = zwK1;
      zp[idx] = zpK;
      if (copy_in) {
        zwi[idx + P] = (T)zwK1;
        zpi[idx] = (T)zpK;
      }
      const double w0 = F3(f.weto, I, J, K);
      if (w0 != 0.0 && w0 != 1.0) nbad = 1;
      unsigned fl = 0;
      if (w0 != 0.0) {
        const double ww = F3(f.weto, I - 1, J, K), we = F3(f.weto, I + 1, J, K);
        const double wn = F3(f.weto, I, J - 1, K), ws = F3(f.weto, I, J + 1, K);
        fl |= 1u;
        if (ww == 1.0 && (f.cyc_x || I != 2)) fl |= 1u << 1;  // couplings, mo_petsc.f90:692-764
        if (we == 1.0 && (f.cyc_x || I != IE1)) fl |= 1u << 2;
        if (wn == 1.0 && (f.cyc_y || J != 2)) fl |= 1u << 3;
        if (ws == 1.0 && (f.cyc_y || J != JE1)) fl |= 1u << 4;
        if (K != 1) fl |= 1u << 5;
        if (K != KE && F3(f.weto, I, J, K + 1) == 1.0) fl |= 1u << 6;
        unsigned xc = 0, yc = 0, zc = 0;  // Neumann at land, mo_petsc.f90:630-653
        if (ww == 0.0 && we == 1.0)
          xc = 1;
        else if (ww == 1.0 && we == 0.0)
          xc = 2;
        else if (ww == 0.0 && we == 0.0)
          xc = 3;
        if (wn == 0.0 && ws == 1.0)
          yc = 1;
        else if (wn == 1.0 && ws == 0.0)
          yc = 2;
        else if (wn == 0.0 && ws == 0.0)
          yc = 3;
        // surface / bottom, mo_petsc.f90:657-665 (K == 1 is tested first)
        if (K == 1)
          zc = 0;
        else if (K == KE || F3(f.weto, I, J, K + 1) == 0.0)
          zc = 1;
        fl |= xc << 7 | yc << 9 | zc << 11;
      }
      flg[idx] = (uint16_t)fl;
      const double Sv = cc[6] / zpK;
      smax = fmax(smax, Sv);
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

// Inner operator row C = S * (-A), S = DLXP*DLYP/ZP, in T arithmetic from the T copies.
template <typename T>
__device__ __forceinline__ void inner_row(const T* cc, T zwk, T zwk1, T zpk, unsigned f, T* c7) {
  row_coef<T>(cc, zwk, zwk1, zpk, f, c7);
  const T S = cc[6] / zpk;
#pragma unroll
  for (int q = 0; q < 7; ++q) c7[q] = S * (-c7[q]);
}

// Line preconditioner factorisation (per column), restated by orc_line_factor*: the tridiagonal of
// the inner operator (centre, and the up coupling of wet-wet pairs as the symmetric off-diagonal).
template <typename T>
__global__ __launch_bounds__(NT) void k_line_factor(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const size_t P = s.plane, col = (size_t)j * s.pitch + i;
  gptr<const T> zwi = G<const T>(s.zw_in);
  gptr<const T> zpi = G<const T>(s.zp_in);
  gptr<const T> cpi = G<const T>(s.col_in);
  gptr<const uint16_t> flg = G<const uint16_t>(s.flg);
  T cc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) cc[q] = cpi[q * P + col];
  // the recurrence of the factorisation; emit(k, invD_k, l_k) receives each level's factors
  auto factor = [&](auto&& emit) {
    T Dprev = T(1), cd_prev = T(0);
    for (int k = 0; k < s.nz; ++k) {
      const size_t n = col + (size_t)k * P;
      T c7[7];
      inner_row<T>(cc, zwi[n], zwi[n + P], zpi[n], flg[n], c7);
      T d = c7[0], l = T(0);
      if (k > 0) {
        const T e = (cd_prev != T(0) && c7[5] != T(0)) ? c7[5] : T(0);
        l = e / Dprev;
        d = fmad(-l, e, d);
      }
      emit(n, T(1) / d, l);
      Dprev = d;
      cd_prev = c7[6];
    }
  };
  if constexpr (std::is_same<T, float>::value) {
    float mx = 0.0f;
    factor([&](size_t, float inv, float) { mx = fmaxf(mx, fabsf(inv)); });
    int ex = 0;
    const bool ok = mx > 0.0f && isfinite(mx);
    if (ok) frexpf(mx, &ex);
    const int sc = ok ? 15 - ex : 0;
    G<float>(s.dscale)[col] = ldexpf(1.0f, -sc);
    gptr<uint16_t> invD = G<uint16_t>(s.invD);
    gptr<uint16_t> L = G<uint16_t>(s.L);
    factor([&](size_t n, float inv, float l) {
      float v = ldexpf(inv, sc);
      if (fabsf(v) < kHalfMin) v = copysignf(kHalfMin, v);
      invD[n] = half_bits(v);
      L[n] = fabsf(l) < kHalfMin ? (uint16_t)0 : half_bits(l);
    });
  } else {
    gptr<T> invD = G<T>(s.invD);
    gptr<T> L = G<T>(s.L);
    factor([&](size_t n, T inv, T l) {
      invD[n] = inv;
      L[n] = l;
    });
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
      const double wet = (s.flg[idx] & 1u) ? 1.0 : 0.0;
      const double v = divg[(size_t)(i + 1) + (size_t)IE * ((size_t)(j + 1) + (size_t)JE * k)] * wet;
      s.b[idx] = v;
      acc = fma(v, v, acc);
    }
  }
  publish_rows<double, TY, TY>(s, t.bx, t.by, acc, dummy);
}

// b <- b * (1 + amplitude * xi), xi ~ N(0,1) per cell from a counter-based SplitMix64 stream
// keyed by (seed of the system, cell index); land cells stay 0. Synthetic time-step forcing for
// serving loads. seeds[q] is system q's seed, so one launch perturbs every system of a batch.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(NT) void k_perturb(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                double amplitude, const unsigned long long* __restrict__ seeds) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  double acc = 0.0, dummy = 0.0;
  if (i < s.nx && j < s.ny) {
    gptr<double> b = G<double>(s.b);
    const size_t col = (size_t)j * s.pitch + i;
    const unsigned long long key = splitmix64(seeds[t.sys]);
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
  publish_rows<double, TY, TY>(s, t.bx, t.by, acc, dummy);
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

__device__ __forceinline__ double apply_a64(const Sys& s, const double* cc, gptr<const double> x, int i, int j, int k,
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
  double c7[7];
  row_coef<double>(cc, G<const double>(s.zw)[idx], G<const double>(s.zw)[idx + P], G<const double>(s.zp)[idx],
                   G<const uint16_t>(s.flg)[idx], c7);
  double acc = c7[0] * x0;
  acc = fma(c7[1], xw, acc);
  acc = fma(c7[2], xe, acc);
  acc = fma(c7[3], xn, acc);
  acc = fma(c7[4], xs, acc);
  acc = fma(c7[5], xu, acc);
  acc = fma(c7[6], xd, acc);
  return acc;
}

__device__ __forceinline__ void load_col(const Sys& s, size_t col, double* cc) {
  gptr<const double> cp = G<const double>(s.col);
#pragma unroll
  for (int q = 0; q < 7; ++q) cc[q] = cp[q * (size_t)s.plane + col];
}

__global__ __launch_bounds__(NT) void k_apply64(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                const double* __restrict__ x, double* __restrict__ y) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const size_t col = (size_t)j * s.pitch + i;
  double cc[7];
  load_col(s, col, cc);
  for (int k = 0; k < s.nz; ++k) {
    const size_t idx = col + (size_t)k * s.plane;
    y[idx] = apply_a64(s, cc, G<const double>(x), i, j, k, idx);
  }
}

// Materialises the 7 planes of the reference operator (verification entry point).
__global__ __launch_bounds__(NT) void k_materialize(const Sys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                    double* __restrict__ A) {
  const Tile t = tiles[blockIdx.x];
  const Sys& s = systems[t.sys];
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * TY + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const size_t P = s.plane, NP = P * s.nz, col = (size_t)j * s.pitch + i;
  double cc[7], c7[7];
  load_col(s, col, cc);
  for (int k = 0; k < s.nz; ++k) {
    const size_t idx = col + (size_t)k * P;
    row_coef<double>(cc, s.zw[idx], s.zw[idx + P], s.zp[idx], s.flg[idx], c7);
#pragma unroll
    for (int q = 0; q < 7; ++q) A[q * NP + idx] = c7[q];
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

// H: rows per tile (blocks of 64 x H threads, one column each; the row sums go to the same 4-row slots
// for every H, so H does not change the bits).
template <typename T, int H>
__global__ __launch_bounds__(64 * H) void k_residual(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (t.sys < 0) return;  // an empty tile of a work-balanced list
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  if (o->odone != RUN) return;
  const double nb = o->nb;
  const bool zero = (nb == 0.0);
  const double gamma = 1.0 / (o->smax * o->nr_prev);
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * H + (threadIdx.x >> 6);
  double acc = 0.0, dummy = 0.0;
  if (i < s.nx && j < s.ny) {
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const size_t col = (size_t)j * s.pitch + i, P = s.plane, IP = s.pitch;
    gptr<T> rin = G<T>(s.r);
    gptr<double> p64w = G<double>(s.p64);
    if (zero) {
      for (int k = 0; k < nz; ++k) {
        p64w[col + (size_t)k * P] = 0.0;
        rin[col + (size_t)k * P] = T(0);
      }
    } else {
      gptr<const double> p = G<const double>(s.p64);
      gptr<const double> b = G<const double>(s.b);
      gptr<const double> zw = G<const double>(s.zw);
      gptr<const double> zp = G<const double>(s.zp);
      gptr<const uint16_t> flg = G<const uint16_t>(s.flg);
      const bool tab = s.tab != 0;
      ZCol<double> zc;
      if (tab)
        zc.load(G<const double>(s.tw), G<const double>(s.tp), s.xk, G<const double>(s.xw), G<const double>(s.xp), s.xf,
                P, col);
      double cc[7];
      load_col(s, col, cc);
      // horizontal neighbour columns (periodic wrap where cyclic; absent ones read this column and
      // are replaced by +0), the same for every level
      const bool hw = i > 0 || s.cyc_x, he = i < nx - 1 || s.cyc_x;
      const bool hn = j > 0 || s.cyc_y, hs = j < ny - 1 || s.cyc_y;
      const size_t cw = i > 0 ? col - 1 : (s.cyc_x ? col + (size_t)(nx - 1) : col);
      const size_t ce = i < nx - 1 ? col + 1 : (s.cyc_x ? col - (size_t)(nx - 1) : col);
      const size_t cn = j > 0 ? col - IP : (s.cyc_y ? col + IP * (size_t)(ny - 1) : col);
      const size_t cs = j < ny - 1 ? col + IP : (s.cyc_y ? col - IP * (size_t)(ny - 1) : col);
      // level data in two register sets used alternately: level k+1 is loaded while level k is
      // computed; p and ZW of the column itself rotate through registers (x(k-1), x(k), ZW(k))
      struct Lv {
        double b, zp, zw1, pd, pw, pe, pn, ps;
        unsigned f;
      };
      auto load = [&](int k, Lv& v) {
        const int kk = k < nz ? k : nz - 1;
        const size_t o = (size_t)kk * P;
        v.b = ld_nt<double>(b, col + o);
        v.zp = tab ? zc.zp(kk) : ld_nt<double>(zp, col + o);
        v.zw1 = tab ? zc.zw(kk + 1) : ld_nt<double>(zw, col + o + P);
        v.f = tab ? zc.flag(kk, nz) : ld_nt<uint16_t>(flg, col + o);
        v.pd = p[col + (kk < nz - 1 ? o + P : o)];
        v.pw = p[cw + o];
        v.pe = p[ce + o];
        v.pn = p[cn + o];
        v.ps = p[cs + o];
      };
      double xu = 0.0, x0 = p[col], zwk = tab ? zc.zw(0) : ld_nt<double>(zw, col);
      auto compute = [&](int k, const Lv& v) {
        const double xd = k < nz - 1 ? v.pd : 0.0;
        double c7[7];
        row_coef<double>(cc, zwk, v.zw1, v.zp, v.f, c7);
        double a = c7[0] * x0;
        a = fma(c7[1], hw ? v.pw : 0.0, a);
        a = fma(c7[2], he ? v.pe : 0.0, a);
        a = fma(c7[3], hn ? v.pn : 0.0, a);
        a = fma(c7[4], hs ? v.ps : 0.0, a);
        a = fma(c7[5], xu, a);
        a = fma(c7[6], xd, a);
        const double r = v.b - a;
        acc = fma(r, r, acc);
        const double S = cc[6] / v.zp;
        st_nt<T>(rin, col + (size_t)k * P, (T)((S * (-r)) * gamma));
        xu = x0;
        x0 = xd;
        zwk = v.zw1;
      };
      Lv l0, l1;
      load(0, l0);
      for (int k = 0; k < nz; k += 2) {
        load(k + 1, l1);
        compute(k, l0);
        if (k + 1 >= nz) break;
        load(k + 2, l0);
        compute(k + 1, l1);
      }
    }
  }
  publish_rows<double, H, H>(s, t.bx, t.by, acc, dummy);
}

template <typename T, int H>
__global__ __launch_bounds__(64 * H) void k_refine(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (t.sys < 0) return;  // an empty tile of a work-balanced list
  const Sys& s = systems[t.sys];
  const Outer* o = s.outer;
  if (o->odone != RUN) return;
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * H + (threadIdx.x >> 6);
  if (i >= s.nx || j >= s.ny) return;
  const double g = o->inv_gamma;
  const Inner<T>* in = reinterpret_cast<const Inner<T>*>(s.inner);
  // the inner solve's last x += alpha p, left pending by its last column update (k_bwd applies
  // the others); p of that iteration is the buffer the next matvec would have read as p_old
  const bool pend = in->xpend != 0;
  const T alpha = in->alpha;
  gptr<const T> x = G<const T>(s.x);
  gptr<const T> pl = G<const T>(in->pbuf ? s.pb : s.pa);
  gptr<double> p64 = G<double>(s.p64);
  const size_t col = (size_t)j * s.pitch + i;
#pragma unroll 4
  for (int k = 0; k < s.nz; ++k) {
    const size_t idx = col + (size_t)k * s.plane;
    const T xv = ld_nt<T>(x, idx);
    const T xk = pend ? fmad(alpha, ld_nt<T>(pl, idx), xv) : xv;
    p64[idx] = p64[idx] + (double)xk * g;
  }
}

// ------------------------------------------------------------------------------------------------
// Inner PCG, two kernels per iteration.
//
//   k_fwd  (per column, top-down):   r -= alpha q; forward elimination of the line preconditioner
//          y_k = r_k - l_k y_{k-1},  w_k = y_k / D_k;  r.z = sum y_k w_k,  r.r.
//          Reads r q l invD, writes r w: 4 words + 2 factors per cell (factors fp16 in FP32 mode).
//   k_bwd  (64 x H tile, bottom-up): back substitution z_k = w_k - l_{k+1} z_{k+1} fused with
//          p = z + beta p_old, q = C p (7-point, horizontal neighbours through LDS) and p.q, and
//          the previous iteration's x += alpha p_old; the rows of C are rebuilt from ZW, ZP, the
//          flag bits and the column's metric products.
//          Reads w l p_old x ZW ZP (+ 2 B of flags), writes p q x: 8 words + factor + 2 B per cell.
// The halo lanes of k_bwd run the back-substitution recurrence of their halo columns themselves, so
// z never goes to memory and no grid-wide step separates the two halves of the line solve.

// U: levels per load batch (the batch after next is loaded while one is computed).
template <typename T, bool INIT, int H, int U>
__global__ __launch_bounds__(64 * H) void k_fwd(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (t.sys < 0) return;  // an empty tile of a work-balanced list
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (INIT) {
    if (o->odone != RUN) return;
  } else {
    if (in->done != RUN) return;
  }
  const int i = t.bx * TX + (threadIdx.x & (TX - 1)), j = t.by * H + (threadIdx.x >> 6);
  const size_t P = s.plane, col = (size_t)j * s.pitch + i;
  const int nz = s.nz;
  gptr<T> x = G<T>(s.x);
  gptr<T> r = G<T>(s.r);
  gptr<T> w = G<T>(s.z);
  gptr<const T> q = G<const T>(s.q);
  gptr<const typename Fac<T>::S> L = G<const typename Fac<T>::S>(s.L);
  gptr<const typename Fac<T>::S> invD = G<const typename Fac<T>::S>(s.invD);
  const T alpha = INIT ? T(0) : in->alpha;
  T rho = T(0), rr = T(0);
  if (i < s.nx && j < s.ny) {
    // FP32 factors: invD is stored scaled by a power of two per column; undoing it is exact
    T cs = T(1);
    if constexpr (std::is_same<T, float>::value) cs = G<const float>(s.dscale)[col];
    // Batches of U levels in two register sets (ping-pong): the loads of batch b+1 are issued before
    // batch b is computed and stored, so stores (which may alias later loads as far as the compiler
    // knows) never sit between a load and its use, and no prefetched value has to be moved.
    struct Batch {
      T r[U], q[U], l[U], d[U];
    };
    Batch b0, b1;
    auto load = [&](int k0, Batch& b) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u < nz ? k0 + u : nz - 1;
        const size_t idx = col + (size_t)k * P;
        b.r[u] = ld_nt<T>(r, idx);
        b.l[u] = Fac<T>::ld_nt(L, idx);
        b.d[u] = Fac<T>::ld_nt(invD, idx) * cs;
        if (!INIT) b.q[u] = ld_nt<T>(q, idx);
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
            st_nt<T>(x, idx, T(0));
          } else {
            rk = fmad(-alpha, b.q[u], rk);
            st_nt<T>(r, idx, rk);
          }
          rr = fmad(rk, rk, rr);
          const T y = fmad(-b.l[u], yprev, rk);
          const T wv = y * b.d[u];
          rho = fmad(y, wv, rho);
          st_nt<T>(w, idx, wv);
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
  publish_rows<T, H, H>(s, t.bx, t.by, rho, rr);
}

// Two adjacent columns per lane: vectors of two values in one load or store.
template <typename T>
struct V2;
template <>
struct V2<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct V2<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
// the two factors of adjacent columns: a pair of fp16 in one 32-bit word (FP32), a vector of 2 (FP64)
template <typename T>
struct Fac2 {
  using S = typename V2<T>::type;
  __device__ static void ld_nt(gptr<const T> p, size_t i, T& a, T& b) {
    const S v = __builtin_nontemporal_load(reinterpret_cast<gptr<const S>>(p + i));
    a = v.x;
    b = v.y;
  }
};
template <>
struct Fac2<float> {
  __device__ static void ld_nt(gptr<const uint16_t> p, size_t i, float& a, float& b) {
    const uint32_t v = __builtin_nontemporal_load(reinterpret_cast<gptr<const uint32_t>>(p + i));
    a = half_val((uint16_t)(v & 0xffffu));
    b = half_val((uint16_t)(v >> 16));
  }
};

// The sum over a tile row held by a half-wavefront, lane h holding columns 2h (v0) and 2h+1 (v1):
// the same tree as wave_sum over one column per lane (column c paired with c+32, c+16, ..., c+1), so
// the row sums are the same bits in either layout.
template <typename T>
__device__ __forceinline__ T half_row_sum(T v0, T v1) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
  }
  return v0 + v1;
}

// publish_rows for a block of two tiles whose wavefront w holds row w of both, the first tile in lanes
// 0-31 and the second in lanes 32-63, with their row sums already formed (half_row_sum): the same
// slots and the same group sums per tile. The second tile is (bx + 1, by), or (bx, by + 1) when vert.
template <typename T, int H>
__device__ __forceinline__ void publish_pair_rows(const Sys& s, int bx, int by, bool vert, T a, T b) {
  static_assert(H % RG == 0, "row groups must not straddle tiles");
  __shared__ T sh[2][2 * H];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, half = l >> 5;
  __syncthreads();
  if ((l & 31) == 0) {
    sh[half][w] = a;
    sh[half][H + w] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2 * (H / RG)) {
    const int half2 = threadIdx.x / (H / RG), g = threadIdx.x % (H / RG);
    const int grow = (by + (vert ? half2 : 0)) * (H / RG) + g, tx = bx + (vert ? 0 : half2);
    if (grow * RG < s.ny && tx < s.tiles_x) {
      const T* v = sh[half2];
      const T sa = (v[4 * g] + v[4 * g + 1]) + (v[4 * g + 2] + v[4 * g + 3]);
      const T sb = (v[H + 4 * g] + v[H + 4 * g + 1]) + (v[H + 4 * g + 2] + v[H + 4 * g + 3]);
      gptr<T> part = G<T>(s.partials);
      const size_t slot = (size_t)grow * s.tiles_x + tx;
      part[2 * slot] = sa;
      part[2 * slot + 1] = sb;
    }
  }
}

// k_fwd over a pair of adjacent 64 x H tiles (bx even, bx+1): wavefront w holds row j0 + w of both,
// lane l the columns 2l and 2l+1 of the pair (the first tile in lanes 0-31), so every load and store
// is one contiguous 512-byte run per wavefront. A block whose tile is marked vertical (Tile::local = 1:
// the last tiles of two row-blocks of a system with an odd number of tiles per row) holds (bx, by) in
// lanes 0-31 and (bx, by + 1) in lanes 32-63, two runs of 256 bytes, instead of two partly filled blocks. The same operations per column, in the same order, and
// the same row sums per tile (half_row_sum), so the same bits as k_fwd; half as many, twice as wide
// memory operations stream the column march faster (1024x768x60: 6.0 against 5.0-5.5 TB/s for a march
// of this shape).
template <typename T, bool INIT, int H, int U>
__global__ __launch_bounds__(64 * H) void k_fwd2(const Sys* __restrict__ systems, const Tile* __restrict__ tiles) {
  using T2 = typename V2<T>::type;
  const Tile t = tiles[blockIdx.x];
  if (t.sys < 0) return;  // an empty tile of a work-balanced list
  const Sys& s = systems[t.sys];
  Outer* o = s.outer;
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (INIT) {
    if (o->odone != RUN) return;
  } else {
    if (in->done != RUN) return;
  }
  const bool vert = t.local != 0;
  const int l = threadIdx.x & 63;
  const int i = t.bx * TX + 2 * (vert ? (l & 31) : l), j = (t.by + (vert ? (l >> 5) : 0)) * H + (threadIdx.x >> 6);
  const size_t P = s.plane, col = (size_t)j * s.pitch + i;
  const int nz = s.nz;
  const bool ok0 = i < s.nx && j < s.ny, ok1 = i + 1 < s.nx && j < s.ny;
  gptr<T2> x = reinterpret_cast<gptr<T2>>(G<T>(s.x));
  gptr<T2> r = reinterpret_cast<gptr<T2>>(G<T>(s.r));
  gptr<T2> w = reinterpret_cast<gptr<T2>>(G<T>(s.z));
  gptr<T> x1 = G<T>(s.x), r1 = G<T>(s.r), w1 = G<T>(s.z);
  gptr<const T2> q = reinterpret_cast<gptr<const T2>>(G<const T>(s.q));
  gptr<const typename Fac<T>::S> L = G<const typename Fac<T>::S>(s.L);
  gptr<const typename Fac<T>::S> invD = G<const typename Fac<T>::S>(s.invD);
  const T alpha = INIT ? T(0) : in->alpha;
  T rho0 = T(0), rr0 = T(0), rho1 = T(0), rr1 = T(0);
  if (ok0) {
    // the pair's element index is col / 2 (i is even and the pitch a multiple of 32); with i + 1
    // beyond nx the second column is padding, read but never stored or summed
    const size_t e2 = col / 2, P2 = P / 2;
    T cs0 = T(1), cs1 = T(1);
    if constexpr (std::is_same<T, float>::value) {
      const T2 c = __builtin_nontemporal_load(reinterpret_cast<gptr<const T2>>(G<const float>(s.dscale)) + e2);
      cs0 = c.x;
      cs1 = c.y;
    }
    struct Batch {
      T2 r[U], q[U];
      T l0[U], l1[U], d0[U], d1[U];
    };
    Batch b0, b1;
    auto load = [&](int k0, Batch& b) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u < nz ? k0 + u : nz - 1;
        const size_t idx = col + (size_t)k * P;
        b.r[u] = __builtin_nontemporal_load(r + (e2 + (size_t)k * P2));
        Fac2<T>::ld_nt(L, idx, b.l0[u], b.l1[u]);
        Fac2<T>::ld_nt(invD, idx, b.d0[u], b.d1[u]);
        b.d0[u] *= cs0;
        b.d1[u] *= cs1;
        if (!INIT) b.q[u] = __builtin_nontemporal_load(q + (e2 + (size_t)k * P2));
      }
    };
    T y0 = T(0), y1 = T(0);
    auto compute = [&](int k0, const Batch& b) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k0 + u < nz) {
          const size_t e = e2 + (size_t)(k0 + u) * P2, idx = col + (size_t)(k0 + u) * P;
          T rk0 = b.r[u].x, rk1 = b.r[u].y;
          if (INIT) {
            if (ok1) {
              __builtin_nontemporal_store(T2{T(0), T(0)}, x + e);
            } else {
              st_nt<T>(x1, idx, T(0));
            }
          } else {
            rk0 = fmad(-alpha, b.q[u].x, rk0);
            rk1 = fmad(-alpha, b.q[u].y, rk1);
            if (ok1) {
              __builtin_nontemporal_store(T2{rk0, rk1}, r + e);
            } else {
              st_nt<T>(r1, idx, rk0);
            }
          }
          rr0 = fmad(rk0, rk0, rr0);
          rr1 = fmad(rk1, rk1, rr1);
          const T ya = fmad(-b.l0[u], y0, rk0), yb = fmad(-b.l1[u], y1, rk1);
          const T wa = ya * b.d0[u], wb = yb * b.d1[u];
          rho0 = fmad(ya, wa, rho0);
          rho1 = fmad(yb, wb, rho1);
          if (ok1) {
            __builtin_nontemporal_store(T2{wa, wb}, w + e);
          } else {
            st_nt<T>(w1, idx, wa);
          }
          y0 = ya;
          y1 = yb;
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
    if (!ok1) {
      rho1 = T(0);
      rr1 = T(0);
    }
  }
  const T rho = half_row_sum(rho0, rho1), rr = half_row_sum(rr0, rr1);
  publish_pair_rows<T, H>(s, t.bx, t.by, vert, rho, rr);
}

// Threads of the matvec kernel (tile of 64 x H columns): H wavefronts own the tile rows (one column
// each, centre slots); three more wavefronts own the halo: the north row, the south row and the 2*H
// west/east cells. Every thread runs the back-substitution recurrence of exactly one column, so no
// wavefront does both centre and halo work between two barriers.
template <int H>
constexpr int bwd_threads() {
  return 64 * (H + 3);
}

// Levels the matvec kernel's loads run ahead of their use (the depth of its software pipeline).
// A launch of many tiles hides the latency of a load behind other blocks' work, and one level ahead
// keeps registers low: 12-row FP64 tiles lose a third of their rate at depth 2 or 4 (occupancy), and
// FP32 tiles run best at depth 1. FP64 tiles of 4 and 8 rows, lighter blocks, gain from 4 levels
// (1024x768x60 FP64 at 8 rows: 19.5 -> 20.5 solves/s). A launch of fewer tiles than CUs
// (kDeepTiles) is a few chains of dependent levels: FP32 prefetches 4 levels there, FP64 runs the
// chunked kernel below, and k_fwd loads longer batches.
constexpr int kDeepDepth = 4;

// With the table form, whether the matvec kernels derive each cell's flag bits from the column's wet
// depths (2 B per cell less) or read the flag plane: FP64 derives (regional FP64 20.6 -> 20.9 solves/s,
// batch 177 -> 181); FP32 reads, because the 6 more registers per thread cost its 12-row blocks their
// second block per CU (mixed 41.3 -> 34.4, or 38.1 when held to 64 registers, with spills; in the
// buffer form, held to 64 registers: 279 against 271 us per 1024x768x60 launch). The chunked FP64
// kernel of 12-row tiles reads too (it would spill).
template <typename T>
constexpr bool kDeriveFlags = sizeof(T) == 8;


// Occupancy the buffer form of the FP32 streaming matvec asks the compiler for: two 12-row blocks
// (15 wavefronts each) per CU need 8 wavefronts per SIMD, four 4-row blocks (7 each) 7; it reaches them
// by spilling a few scalar registers into vector-register lanes. The other forms keep the compiler's
// own choice.
template <typename T, int H, int D, bool BUF>
constexpr int bwd_waves() {
  return (BUF && sizeof(T) == 4 && D == 1) ? (H == 12 ? 8 : H == 4 ? 7 : 1) : 1;
}

// raw != 0 (verification only): w is taken as z itself (l treated as 0), so q = C w.
// BUF: the column march addresses memory through buffer descriptors (every array of the launch at most
// kBufMaxBytes); else through 64-bit pointers with branches around the masked accesses. Same bits.
template <typename T, int H, int D, bool BUF>
__global__ __launch_bounds__(bwd_threads<H>()) __attribute__((amdgpu_waves_per_eu(bwd_waves<T, H, D, BUF>())))
void k_bwd(const Sys* __restrict__ systems, const Tile* __restrict__ tiles, int raw) {
  const Tile t = tiles[blockIdx.x];
  if (t.sys < 0) return;  // an empty tile of a work-balanced list
  const Sys& s = systems[t.sys];
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (in->done != RUN) return;
  // the wavefront's row and the inner state's flags as wave-uniform values (the inner state is read
  // with vector loads), so that the masks and descriptors derived from them stay scalar
  const int tx = threadIdx.x & 63, ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = t.bx * TX, j0 = t.by * H;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const size_t P = s.plane, IP = s.pitch;
  const bool first = __builtin_amdgcn_readfirstlane(in->first) != 0;
  const int pb = __builtin_amdgcn_readfirstlane(in->pbuf);
  const T beta = in->beta;
  // the previous iteration's solution update x += alpha p_old, deferred to here where p_old is read
  // anyway (not on an inner solve's first iteration, not in the verification mode)
  const bool xupd = !first && !raw;
  const T alpha = in->alpha;
  using FS = typename Fac<T>::S;

  __shared__ T lds[3][H + 2][TX + 2];

  // this thread's slot in the LDS plane and the column it follows
  int gi, gj, sx, sy;
  bool slot = true;
  const bool centre = ty < H;
  if (centre) {
    gi = i0 + tx;
    gj = j0 + ty;
    sx = tx + 1;
    sy = ty + 1;
  } else if (ty == H) {
    gi = i0 + tx;
    gj = j0 - 1;
    sx = tx + 1;
    sy = 0;
  } else if (ty == H + 1) {
    gi = i0 + tx;
    gj = j0 + H;
    sx = tx + 1;
    sy = H + 1;
  } else {
    slot = tx < 2 * H;
    const int e = tx < H ? 0 : 1, r = tx < H ? tx : tx - H;
    gi = e ? i0 + TX : i0 - 1;
    gj = j0 + (slot ? r : 0);
    sx = e ? TX + 1 : 0;
    sy = (slot ? r : 0) + 1;
  }
  const bool active = centre && gi < nx && gj < ny;
  int mi = gi, mj = gj;
  const bool val = slot && map_ij(mi, mj, nx, ny, s.cyc_x, s.cyc_y);
  const size_t off = val ? (size_t)mj * IP + mi : 0;

  // Software pipeline over levels (bottom-up) in D+1 register sets used in turn: while level k is
  // finished with set `cur` (p at level k-1 from its w/l/p_old, q at level k from its row data), set
  // `nxt` receives w/l/p_old of level k-1-D and the row data (ZW, ZP, flags) of level k-D, the set
  // that step k-D will finish with. The 7 inner coefficients are rebuilt from the row data and this
  // column's metric products. A set holds the loaded values as loaded (fp16 factor bits, the plane's
  // or the table's ZW/ZP): they are decoded and selected where they are used, so that no load is
  // waited for before its level.
  struct Stage {
    T w, p, x;
    FS l;
    T zw, zp, tw, tp;
    unsigned f;
  };
  T cc[7];
  const bool tab = s.tab != 0;
  const bool derive = tab && kDeriveFlags<T>;
  ZCol<T> zcol;
  if (centre) {
    gptr<const T> cpi = G<const T>(s.col_in);
#pragma unroll
    for (int c = 0; c < 7; ++c) cc[c] = cpi[c * P + off];
    if (tab)
      zcol.load(G<const T>(s.tw_in), G<const T>(s.tp_in), s.xk, G<const T>(s.xw_in), G<const T>(s.xp_in), s.xf, P,
                off);
  }
  T zw_below = T(0);  // ZW at the bottom face of the current level (top face of the level below)
  T zc = T(0), lc = T(0);  // z and l of this thread's column at the level below

  // Memory of the march. Loads: w, l, p_old of every slot lane's column (zero-filled outside the
  // domain), x of the centre lanes in the domain when the update is due, the row data of the centre
  // lanes (planes without the table; with it the table's entry, the same for the whole block, by a
  // scalar load). Stores: p, x and q of the centre lanes in the domain. Level m < 0 (past the surface,
  // the last step's level) stores nothing.
  const void* pold_base = pb ? s.pb : s.pa;
  void* pnew_base = pb ? s.pa : s.pb;
  // buffer form: descriptors and lane offsets
  const uint32_t PBu = (uint32_t)(P * sizeof(T)), PFu = (uint32_t)(P * sizeof(FS)), P2u = (uint32_t)(P * 2);
  const uint32_t o_col = val ? (uint32_t)(off * sizeof(T)) : kOob;
  const uint32_t o_fac = val ? (uint32_t)(off * sizeof(FS)) : kOob;
  const uint32_t o_h = val ? (uint32_t)(off * 2) : kOob;
  const uint32_t o_st = active ? (uint32_t)(off * sizeof(T)) : kOob;
  const uint32_t AB = PBu * (uint32_t)nz, AF = PFu * (uint32_t)nz, A2 = P2u * (uint32_t)nz;
  const uint32_t rec_c = centre ? AB : 0u, rec_x = (centre && xupd) ? AB : 0u;
  const Rsrc r_w = buf_of(s.z, AB), r_l = buf_of(s.L, AF), r_po = buf_of(pold_base, AB);
  const Rsrc r_x = buf_of(s.x, rec_x), r_pn = buf_of(pnew_base, rec_c), r_q = buf_of(s.q, rec_c);
  const Rsrc r_zw = buf_of(s.zw_in, tab ? 0u : rec_c), r_zp = buf_of(s.zp_in, tab ? 0u : rec_c);
  const Rsrc r_f = buf_of(s.flg, (centre && !derive) ? A2 : 0u);
  typedef __attribute__((address_space(4))) const T* cptr;  // read-only during the kernel: scalar loads
  const cptr ctw = (cptr)s.tw_in, ctp = (cptr)s.tp_in;
  // pointer form
  gptr<const T> w = G<const T>(s.z);
  gptr<const FS> L = G<const FS>(s.L);
  gptr<T> xv = G<T>(s.x);
  gptr<const T> pold = G<const T>(pold_base);
  gptr<T> pnew = G<T>(pnew_base);
  gptr<T> q = G<T>(s.q);
  gptr<const T> zwi = G<const T>(s.zw_in);
  gptr<const T> zpi = G<const T>(s.zp_in);
  gptr<const uint16_t> flg = G<const uint16_t>(s.flg);

  auto load_a = [&](int m, Stage& st) {
    const uint32_t mm = (uint32_t)(m > 0 ? m : 0);
    if constexpr (BUF) {
      st.w = bld<T, kAuxNone>(r_w, o_col + mm * PBu);
      st.l = bld<FS, kAuxNone>(r_l, o_fac + mm * PFu);
      st.p = bld<T, kAuxNone>(r_po, o_col + mm * PBu);
      st.x = bld<T, kAuxNt>(r_x, o_st + mm * PBu);
    } else {
      const size_t om = (size_t)mm * P + off;
      st.w = w[om];
      st.l = L[om];
      st.p = pold[om];
      if (centre && xupd) st.x = ld_nt<T>(xv, om);
    }
  };
  auto load_c = [&](int m, Stage& st) {
    const uint32_t mm = (uint32_t)(m > 0 ? m : 0);
    if constexpr (BUF) {
      st.zw = bld<T, kAuxNt>(r_zw, o_col + mm * PBu);
      st.zp = bld<T, kAuxNt>(r_zp, o_col + mm * PBu);
      st.f = bld<uint16_t, kAuxNt>(r_f, o_h + mm * P2u);
      st.tw = tab ? ctw[mm] : T(0);
      st.tp = tab ? ctp[mm] : T(0);
    } else {
      if (!centre) return;
      const size_t cm = (size_t)mm * P + off;
      if (tab) {
        st.zw = zcol.zw(mm);
        st.zp = zcol.zp(mm);
      } else {
        st.zw = ld_nt<T>(zwi, cm);
        st.zp = ld_nt<T>(zpi, cm);
      }
      st.f = derive ? zcol.flag(mm, nz) : ld_nt<uint16_t>(flg, cm);
    }
  };
  auto store = [&](Rsrc r, gptr<T> a, int m, T v) {
    if constexpr (BUF) {
      bst_nt<T>(r, m >= 0 ? o_st + (uint32_t)m * PBu : kOob, v);
    } else {
      if (active && m >= 0) st_nt<T>(a, (size_t)m * P + off, v);
    }
  };
  // back substitution and search-direction update of level m for this thread's column
  auto level = [&](int m, const Stage& st) -> T {
    const T zm = fmad(-lc, zc, st.w);
    const T pm = first ? zm : fmad(beta, st.p, zm);
    zc = zm;
    lc = raw ? T(0) : Fac<T>::dec(st.l);
    const T pout = val ? pm : T(0);
    if (slot) lds[(m + 3) % 3][sy][sx] = pout;
    store(r_pn, pnew, m, pm);
    if (BUF || xupd) store(r_x, xv, m, fmad(alpha, st.p, st.x));
    return pout;
  };
  T p0 = T(0), pp1 = T(0), dot = T(0);
  // step k: row k with p of levels k+1, k, k-1; level k-1 is advanced first. On the last step (k = 0)
  // the buffer form runs the level past the surface with its stores masked and takes its p as 0, so
  // that the step's memory instructions are those of every other step; the pointer form skips it.
  auto step = [&](int k, Stage& cur, Stage& nxt) {
    load_a(k - 1 - D, nxt);
    load_c(k - D, nxt);
    T pm1 = T(0);
    if (BUF) {
      const T pl = level(k - 1, cur);
      pm1 = k > 0 ? pl : T(0);
    } else if (k > 0) {
      pm1 = level(k - 1, cur);
    }
    __syncthreads();
    T acc = T(0);
    if (centre) {
      const int b0 = k % 3;
      const T zwk = (BUF && tab) ? zcol.zw_or(k, cur.tw) : cur.zw;
      const T zpk = (BUF && tab) ? zcol.zp_or(k, cur.tp) : cur.zp;
      const unsigned fk = (BUF && derive) ? zcol.flag(k, nz) : cur.f;
      T c7[7];
      inner_row<T>(cc, zwk, zw_below, zpk, fk, c7);
      zw_below = zwk;
      acc = c7[0] * p0;
      acc = fmad(c7[1], lds[b0][sy][sx - 1], acc);
      acc = fmad(c7[2], lds[b0][sy][sx + 1], acc);
      acc = fmad(c7[3], lds[b0][sy - 1][sx], acc);
      acc = fmad(c7[4], lds[b0][sy + 1][sx], acc);
      acc = fmad(c7[5], pm1, acc);
      acc = fmad(c7[6], pp1, acc);
      if (active) dot = fmad(p0, acc, dot);
    }
    store(r_q, q, k, acc);
    pp1 = p0;
    p0 = pm1;
  };
  if (centre) zw_below = tab ? zcol.zw(nz) : zwi[(size_t)nz * P + off];  // ZW(KE+1), the last level's bottom face
  if constexpr (D == 1) {
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
  } else {
    // set r finishes the steps k = nz-1-r (mod D+1); set D first carries w/l/p_old of the bottom level
    Stage S[D + 1];
    load_a(nz - 1, S[D]);
#pragma unroll
    for (int r = 0; r < D; ++r) {
      load_a(nz - 2 - r, S[r]);
      load_c(nz - 1 - r, S[r]);
    }
    p0 = level(nz - 1, S[D]);
    __syncthreads();
    for (int kb = nz - 1; kb >= 0; kb -= D + 1) {
#pragma unroll
      for (int r = 0; r <= D; ++r) {
        if (kb - r < 0) break;
        step(kb - r, S[r], S[(r + D) % (D + 1)]);
      }
    }
  }
  publish_rows<T, H, H + 3>(s, t.bx, t.by, dot, T(0));
}

// The FP64 matvec kernel for launches of fewer tiles than CUs, where each block's walk down its columns
// is the critical path. The operations are k_bwd's, in k_bwd's order per column (same bits), but the
// levels go in chunks of CH: each thread first runs the back substitution and search-direction update
// of its column through the chunk -- one dependent fma per level, the chunk's loads issued together --
// and writes p to an LDS plane per level; after one barrier the rows of the chunk's levels, which no
// longer depend on one another, are applied. One barrier per chunk instead of one per level.
// The LDS ring holds CH+1 planes (a power of two): a chunk's rows read the planes of its own levels
// and the one above them; p of the level below comes from the row before, in a register.
template <typename T, int H, int CH>
__global__ __launch_bounds__(bwd_threads<H>()) void k_bwd_ch(const Sys* __restrict__ systems,
                                                   const Tile* __restrict__ tiles, int raw) {
  static_assert(((CH + 1) & CH) == 0, "CH + 1 must be a power of two");
  constexpr int NS = CH + 1;
  const Tile t = tiles[blockIdx.x];
  if (t.sys < 0) return;  // an empty tile of a work-balanced list
  const Sys& s = systems[t.sys];
  Inner<T>* in = reinterpret_cast<Inner<T>*>(s.inner);
  if (in->done != RUN) return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i0 = t.bx * TX, j0 = t.by * H;
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const size_t P = s.plane, IP = s.pitch;
  const bool first = in->first != 0;
  const int pb = in->pbuf;
  const T beta = in->beta;
  const bool xupd = !first && !raw;
  const T alpha = in->alpha;
  gptr<const T> w = G<const T>(s.z);
  gptr<const typename Fac<T>::S> L = G<const typename Fac<T>::S>(s.L);
  gptr<T> xv = G<T>(s.x);
  gptr<const T> pold = G<const T>(pb ? s.pb : s.pa);
  gptr<T> pnew = G<T>(pb ? s.pa : s.pb);
  gptr<T> q = G<T>(s.q);
  gptr<const T> zwi = G<const T>(s.zw_in);
  gptr<const T> zpi = G<const T>(s.zp_in);
  gptr<const uint16_t> flg = G<const uint16_t>(s.flg);

  __shared__ T lds[NS][H + 2][TX + 2];

  // this thread's slot in the LDS plane and the column it follows (as k_bwd)
  int gi, gj, sx, sy;
  bool slot = true;
  const bool centre = ty < H;
  if (centre) {
    gi = i0 + tx;
    gj = j0 + ty;
    sx = tx + 1;
    sy = ty + 1;
  } else if (ty == H) {
    gi = i0 + tx;
    gj = j0 - 1;
    sx = tx + 1;
    sy = 0;
  } else if (ty == H + 1) {
    gi = i0 + tx;
    gj = j0 + H;
    sx = tx + 1;
    sy = H + 1;
  } else {
    slot = tx < 2 * H;
    const int e = tx < H ? 0 : 1, r = tx < H ? tx : tx - H;
    gi = e ? i0 + TX : i0 - 1;
    gj = j0 + (slot ? r : 0);
    sx = e ? TX + 1 : 0;
    sy = (slot ? r : 0) + 1;
  }
  const bool active = centre && gi < nx && gj < ny;
  int mi = gi, mj = gj;
  const bool val = slot && map_ij(mi, mj, nx, ny, s.cyc_x, s.cyc_y);
  const size_t off = val ? (size_t)mj * IP + mi : 0;

  T cc[7];
  const bool tab = s.tab != 0;
  ZCol<T> zcol;
  if (centre) {
    gptr<const T> cpi = G<const T>(s.col_in);
#pragma unroll
    for (int c = 0; c < 7; ++c) cc[c] = cpi[c * P + off];
    if (tab)
      zcol.load(G<const T>(s.tw_in), G<const T>(s.tp_in), s.xk, G<const T>(s.xw_in), G<const T>(s.xp_in), s.xf, P,
                off);
  }
  T zw_below = T(0);
  if (centre) zw_below = tab ? zcol.zw(nz) : zwi[(size_t)nz * P + off];  // ZW(KE+1), the last level's bottom face
  T zc = T(0), lc = T(0);  // z and l of this thread's column at the level below

  // back substitution and search-direction update of level m (k_bwd's `level`)
  auto level = [&](int m, T wv, T lv, T pv, T xv_old) {
    const T zm = fmad(-lc, zc, wv);
    const T pm = first ? zm : fmad(beta, pv, zm);
    zc = zm;
    lc = raw ? T(0) : lv;
    if (slot) lds[m & (NS - 1)][sy][sx] = val ? pm : T(0);
    if (active) {
      st_nt<T>(pnew, (size_t)m * P + off, pm);
      if (xupd) st_nt<T>(xv, (size_t)m * P + off, fmad(alpha, pv, xv_old));
    }
  };
  auto ld_x = [&](size_t om) { return (centre && xupd) ? ld_nt<T>(xv, om) : T(0); };
  {
    const size_t om = (size_t)(nz - 1) * P + off;
    level(nz - 1, w[om], Fac<T>::ld(L, om), pold[om], ld_x(om));
  }
  T pp1 = T(0), dot = T(0);
  T aw[CH], al[CH], ap[CH], ax[CH], rzw[CH], rzp[CH];
  unsigned rf[CH];
  for (int kh = nz - 1; kh >= 0; kh -= CH) {
    const int n = kh + 1 < CH ? kh + 1 : CH;  // rows kh .. kh-n+1; back substitution of levels kh-1 .. kh-n
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (u < n) {
        const int m = kh - 1 - u;
        const size_t om = (size_t)(m > 0 ? m : 0) * P + off;
        aw[u] = w[om];
        al[u] = Fac<T>::ld(L, om);
        ap[u] = pold[om];
        ax[u] = ld_x(om);
        if (centre) {
          const int k = kh - u;
          const size_t ck = (size_t)k * P + off;
          rzw[u] = tab ? zcol.zw(k) : ld_nt<T>(zwi, ck);
          rzp[u] = tab ? zcol.zp(k) : ld_nt<T>(zpi, ck);
          rf[u] = (tab && kDeriveFlags<T> && H < 12) ? zcol.flag(k, nz) : ld_nt<uint16_t>(flg, ck);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u)
      if (u < n && kh - 1 - u >= 0) level(kh - 1 - u, aw[u], al[u], ap[u], ax[u]);
    __syncthreads();
    if (centre) {
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        if (u < n) {
          const int k = kh - u, b0 = k & (NS - 1);
          const T p0 = lds[b0][sy][sx];
          const T pm1 = k > 0 ? lds[(k - 1) & (NS - 1)][sy][sx] : T(0);
          T c7[7];
          inner_row<T>(cc, rzw[u], zw_below, rzp[u], rf[u], c7);
          zw_below = rzw[u];
          T acc = c7[0] * p0;
          acc = fmad(c7[1], lds[b0][sy][sx - 1], acc);
          acc = fmad(c7[2], lds[b0][sy][sx + 1], acc);
          acc = fmad(c7[3], lds[b0][sy - 1][sx], acc);
          acc = fmad(c7[4], lds[b0][sy + 1][sx], acc);
          acc = fmad(c7[5], pm1, acc);
          acc = fmad(c7[6], pp1, acc);
          if (active) {
            st_nt<T>(q, (size_t)k * P + off, acc);
            dot = fmad(p0, acc, dot);
          }
          pp1 = p0;
        }
      }
    }
    __syncthreads();
  }
  publish_rows<T, H, H + 3>(s, t.bx, t.by, dot, T(0));
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
  for (int t = threadIdx.x; t < s.nslots; t += NT) {
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
    in->xpend = 0;
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
    in->xpend = 1;  // x += alpha p of this iteration: applied by the next matvec or by k_refine
    in->done = d;
  } else {  // RED_BWD
    const T pq = a;
    const T alpha = in->rho / pq;
    in->pq = pq;
    in->alpha = alpha;
    in->xpend = 0;  // the matvec just applied the previous iteration's update
    if (!(pq > T(0)) || !isfinite((double)alpha)) {
      in->done = BREAKDOWN;
      if (in->iter == 0) o->odone = BREAKDOWN;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Plain copy, one 16 B element per lane and a grid that covers the array: blocks are dispatched in
// address order, so the accesses in flight form one sliding window (a grid-stride loop spreads them
// over the whole array and reaches about 4.7-5.0 TB/s instead of about 6.2). The device's
// achievable streaming bandwidth, and the known-size reference for the HBM byte counters.

__global__ __launch_bounds__(256) void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n4) b[i] = a[i];
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

hipError_t launch_zwindow(hipStream_t st, bool fp32, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                          const FortranIn& f, int* d_bad) {
  if (ntiles <= 0) return hipSuccess;
  if (fp32) hipLaunchKernelGGL(k_ztab<float>, dim3(1), dim3(64), 0, st, d_sys, f);
  else hipLaunchKernelGGL(k_ztab<double>, dim3(1), dim3(64), 0, st, d_sys, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (fp32)
    hipLaunchKernelGGL(k_zwindow<float>, dim3(ntiles), dim3(NT), 0, st, d_sys, d_tiles, d_bad);
  else
    hipLaunchKernelGGL(k_zwindow<double>, dim3(ntiles), dim3(NT), 0, st, d_sys, d_tiles, d_bad);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_zflags, dim3(ntiles), dim3(NT), 0, st, d_sys, d_tiles, f, d_bad);
  return hipGetLastError();
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
                          double amplitude, const unsigned long long* d_seeds) {
  PHASE(k_perturb, (launch_reduce<double, RED_NB>(st, d_sys, nsys)), ntiles, st, d_sys, d_tiles, amplitude, d_seeds);
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

template <typename T, int H>
static hipError_t residual_h(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int nsys) {
  if (ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_residual<T, H>), dim3(ntiles), dim3(64 * H), 0, st, d_sys, d_tiles);
  const hipError_t e = hipGetLastError();
  return e != hipSuccess ? e : launch_reduce<T, RED_RES>(st, d_sys, nsys);
}

template <typename T, int H>
static hipError_t refine_h(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_refine<T, H>), dim3(ntiles), dim3(64 * H), 0, st, d_sys, d_tiles);
  return hipGetLastError();
}

hipError_t launch_residual(hipStream_t st, bool fp32, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                           int nsys) {
  if (h == kOuterRows)
    return fp32 ? residual_h<float, kOuterRows>(st, d_sys, d_tiles, ntiles, nsys)
                : residual_h<double, kOuterRows>(st, d_sys, d_tiles, ntiles, nsys);
  if (h == TY)
    return fp32 ? residual_h<float, TY>(st, d_sys, d_tiles, ntiles, nsys)
                : residual_h<double, TY>(st, d_sys, d_tiles, ntiles, nsys);
  return hipErrorInvalidValue;
}

hipError_t launch_refine(hipStream_t st, bool fp32, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  if (h == kOuterRows)
    return fp32 ? refine_h<float, kOuterRows>(st, d_sys, d_tiles, ntiles)
                : refine_h<double, kOuterRows>(st, d_sys, d_tiles, ntiles);
  if (h == TY)
    return fp32 ? refine_h<float, TY>(st, d_sys, d_tiles, ntiles) : refine_h<double, TY>(st, d_sys, d_tiles, ntiles);
  return hipErrorInvalidValue;
}

template <typename T, int H, int U>
static hipError_t fwd_hu2(hipStream_t st, bool init, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  const dim3 grid(ntiles), block(64 * H);
  if (init)
    hipLaunchKernelGGL((k_fwd2<T, true, H, U>), grid, block, 0, st, d_sys, d_tiles);
  else
    hipLaunchKernelGGL((k_fwd2<T, false, H, U>), grid, block, 0, st, d_sys, d_tiles);
  return hipGetLastError();
}

template <typename T, int H, int U>
static hipError_t fwd_hu(hipStream_t st, bool init, const Sys* d_sys, const Tile* d_tiles, int ntiles) {
  const dim3 grid(ntiles), block(64 * H);
  if (init)
    hipLaunchKernelGGL((k_fwd<T, true, H, U>), grid, block, 0, st, d_sys, d_tiles);
  else
    hipLaunchKernelGGL((k_fwd<T, false, H, U>), grid, block, 0, st, d_sys, d_tiles);
  return hipGetLastError();
}

// Load batches of 4 levels keep many blocks' loads in flight at a time; a launch of fewer tiles than
// CUs is a few column chains, and longer batches keep each chain's loads a memory round trip ahead of
// their use: 16 levels (FP32, 4-row tiles) or 8, as many as fit the registers without spilling.
template <int H>
static hipError_t fwd_h(hipStream_t st, bool fp32, bool init, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                        const Tile* d_pairs, int npairs) {
  constexpr int UF = H < 8 ? 16 : 8, UD = H < 12 ? 8 : 4;
  const bool deep = ntiles < kDeepTiles;
  // load batches of the paired form. FP32 loads one level at a time: 46 registers keep 8 waves per SIMD,
  // where two levels (74 registers, 6 waves) streamed 2% slower at 1024x768x60.
  constexpr int PF = 1, PD = H < 12 ? 4 : 2;
  // A system with an odd number of 64-column tiles leaves its last tile without a partner, and half
  // the lanes of that block idle. Running those tiles one column per lane in a launch of their own
  // (same bits) left the 64-system batch's rate unchanged (362.9 against 362.8 solves/s): a streaming
  // launch is bandwidth-bound, and a half-empty block only streams fewer bytes.
  if (fp32)
    return deep ? fwd_hu<float, H, UF>(st, init, d_sys, d_tiles, ntiles)
                : fwd_hu2<float, H, PF>(st, init, d_sys, d_pairs, npairs);
  return deep ? fwd_hu<double, H, UD>(st, init, d_sys, d_tiles, ntiles)
              : fwd_hu2<double, H, PD>(st, init, d_sys, d_pairs, npairs);
}

template <typename T, int H, int D>
static hipError_t bwd_stream(hipStream_t st, bool buf, const Sys* d_sys, const Tile* d_tiles, int ntiles, int raw) {
  if (buf)
    hipLaunchKernelGGL((k_bwd<T, H, D, true>), dim3(ntiles), dim3(bwd_threads<H>()), 0, st, d_sys, d_tiles, raw);
  else
    hipLaunchKernelGGL((k_bwd<T, H, D, false>), dim3(ntiles), dim3(bwd_threads<H>()), 0, st, d_sys, d_tiles, raw);
  return hipGetLastError();
}

template <typename T, int H, int CH>
static hipError_t bwd_chunked(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, int raw) {
  hipLaunchKernelGGL((k_bwd_ch<T, H, CH>), dim3(ntiles), dim3(bwd_threads<H>()), 0, st, d_sys, d_tiles, raw);
  return hipGetLastError();
}

// Which matvec kernel a launch runs (all give the same bits), as measured on gfx950:
//   FP32, 256 tiles or more: k_bwd, 1 level ahead; fewer: k_bwd, 4 levels ahead (512x4x100: 62 -> 58 us;
//         the chunked form took 76-87 us there for every chunk length tried, 3 to 15)
//   FP64, 12-row tiles: k_bwd, 1 level ahead; 4 and 8 rows: 4 levels ahead; fewer than 256 tiles:
//         k_bwd_ch in chunks of 7 levels (4 rows) or 3 (8 and 12 rows, the most that fit the registers
//         without spilling) (512x4x100: 92 -> 86 us; 128x128x30: 38 -> 34 us)
//   Each k_bwd form addresses memory through buffer descriptors when the context allows it (buf: every
//   inner array of the launch at most kBufMaxBytes), else through 64-bit pointers (same bits). The
//   descriptors took the 1024x768x60 mixed matvec from 282 to 271 us, the FP64 one at 12 rows from 586
//   to 530 us, and the 512x4x100 mixed solve from 25.6 to 27.6 solves/s.
template <int H>
static hipError_t bwd_h(hipStream_t st, bool fp32, bool buf, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                        int raw) {
  const bool few = ntiles < kDeepTiles;
  if (fp32)
    return few ? bwd_stream<float, H, kDeepDepth>(st, buf, d_sys, d_tiles, ntiles, raw)
               : bwd_stream<float, H, 1>(st, buf, d_sys, d_tiles, ntiles, raw);
  if (few) return bwd_chunked<double, H, (H < 8 ? 7 : 3)>(st, d_sys, d_tiles, ntiles, raw);
  if constexpr (H < 12)
    return bwd_stream<double, H, kDeepDepth>(st, buf, d_sys, d_tiles, ntiles, raw);
  else
    return bwd_stream<double, H, 1>(st, buf, d_sys, d_tiles, ntiles, raw);
}

hipError_t launch_fwd(hipStream_t st, bool fp32, bool init, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                      const Tile* d_pairs, int npairs) {
  if (ntiles <= 0) return hipSuccess;
  switch (h) {
    case 4: return fwd_h<4>(st, fp32, init, d_sys, d_tiles, ntiles, d_pairs, npairs);
    case 8: return fwd_h<8>(st, fp32, init, d_sys, d_tiles, ntiles, d_pairs, npairs);
    case 12: return fwd_h<12>(st, fp32, init, d_sys, d_tiles, ntiles, d_pairs, npairs);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_bwd(hipStream_t st, bool fp32, bool buf, int h, const Sys* d_sys, const Tile* d_tiles, int ntiles,
                      int raw) {
  if (ntiles <= 0) return hipSuccess;
  switch (h) {
    case 4: return bwd_h<4>(st, fp32, buf, d_sys, d_tiles, ntiles, raw);
    case 8: return bwd_h<8>(st, fp32, buf, d_sys, d_tiles, ntiles, raw);
    case 12: return bwd_h<12>(st, fp32, buf, d_sys, d_tiles, ntiles, raw);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_inner_reduce(hipStream_t st, bool fp32, InnerPhase ph, const Sys* d_sys, int nsys) {
  switch (ph) {
    case PH_INIT:
      return fp32 ? launch_reduce<float, RED_INIT>(st, d_sys, nsys) : launch_reduce<double, RED_INIT>(st, d_sys, nsys);
    case PH_BWD:
      return fp32 ? launch_reduce<float, RED_BWD>(st, d_sys, nsys) : launch_reduce<double, RED_BWD>(st, d_sys, nsys);
    case PH_FWD:
      return fp32 ? launch_reduce<float, RED_FWD>(st, d_sys, nsys) : launch_reduce<double, RED_FWD>(st, d_sys, nsys);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_copy4(hipStream_t st, const void* src, void* dst, size_t bytes) {
  const size_t n4 = bytes / 16;
  if (n4 == 0) return hipSuccess;
  const size_t blocks = (n4 + 255) / 256;
  if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy4, dim3((unsigned)blocks), dim3(256), 0, st, (const float4*)src, (float4*)dst, n4);
  return hipGetLastError();
}

hipError_t launch_materialize(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, double* A) {
  LAUNCH(k_materialize, ntiles, st, d_sys, d_tiles, A);
}

hipError_t launch_apply64(hipStream_t st, const Sys* d_sys, const Tile* d_tiles, int ntiles, const double* x,
                          double* y) {
  LAUNCH(k_apply64, ntiles, st, d_sys, d_tiles, x, y);
}

}  // namespace onh
