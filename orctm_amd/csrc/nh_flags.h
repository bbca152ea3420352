// Flag bits of the compact operator derived from wet depths: on a z-level grid (WETO wet from the
// surface down to a column's depth, boden.f90:182-213) the flag bits k_assemble sets per cell (the
// branches of mo_petsc.f90:623-765) follow from the wet depths of the cell's column and of its W, E,
// N, S neighbours and from whether each coupling can exist at all (periodic, or not on the domain
// edge). Plain C++ for the host, host+device under hipcc.
#pragma once
#include <cstddef>
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define ONH_HD __host__ __device__ __forceinline__
#else
#define ONH_HD inline
#endif

namespace onh {

// Level K (1-based) of KE; kc, kw, ke, kn, ks: wet depths; edge bits 1, 2, 4, 8: the W, E, N, S
// coupling can exist. Bit 0 wet; bits 1-6 couplings W, E, N, S, U, D; bits 7-8 / 9-10 the X_C / Y_C
// variant (1: west / north land, 2: east / south land, 3: both); bit 11 the bottom Z_C.
ONH_HD unsigned derive_flags(int K, int KE, int kc, int kw, int ke, int kn, int ks, unsigned edge) {
  if (K > kc) return 0u;
  const bool ww = K <= kw, we = K <= ke, wn = K <= kn, ws = K <= ks, below = K < kc && K != KE;
  unsigned fl = 1u;
  if (ww && (edge & 1u)) fl |= 1u << 1;
  if (we && (edge & 2u)) fl |= 1u << 2;
  if (wn && (edge & 4u)) fl |= 1u << 3;
  if (ws && (edge & 8u)) fl |= 1u << 4;
  if (K != 1) fl |= 1u << 5;
  if (below) fl |= 1u << 6;
  const unsigned xc = !ww ? (we ? 1u : 3u) : (we ? 0u : 2u);
  const unsigned yc = !wn ? (ws ? 1u : 3u) : (ws ? 0u : 2u);
  const unsigned zc = (K != 1 && !below) ? 1u : 0u;
  return fl | xc << 7 | yc << 9 | zc << 11;
}

// The 3 words of interior column (I, J) (1-based, halo'd Fortran indices) from WETO(IE,JE,KE):
// wet depth of the column | W neighbour << 16; E | N << 16; S | edge bits << 16.
ONH_HD void column_flag_words(const double* weto, int IE, int JE, int KE, int I, int J, int cyc_x, int cyc_y,
                              uint32_t* w) {
  auto depth = [&](int Ic, int Jc) {
    int d = 0;
    while (d < KE && weto[(Ic - 1) + (std::size_t)IE * ((Jc - 1) + (std::size_t)JE * d)] == 1.0) ++d;
    return d;
  };
  const unsigned edge = ((cyc_x || I != 2) ? 1u : 0u) | ((cyc_x || I != IE - 1) ? 2u : 0u) |
                        ((cyc_y || J != 2) ? 4u : 0u) | ((cyc_y || J != JE - 1) ? 8u : 0u);
  w[0] = (uint32_t)depth(I, J) | (uint32_t)depth(I - 1, J) << 16;
  w[1] = (uint32_t)depth(I + 1, J) | (uint32_t)depth(I, J - 1) << 16;
  w[2] = (uint32_t)depth(I, J + 1) | edge << 16;
}

ONH_HD unsigned flags_from_words(int K, int KE, const uint32_t* w) {
  return derive_flags(K, KE, (int)(w[0] & 0xffffu), (int)(w[0] >> 16), (int)(w[1] & 0xffffu), (int)(w[1] >> 16),
                      (int)(w[2] & 0xffffu), w[2] >> 16);
}

}  // namespace onh
