// nh_ocv.hip -- the steps of OCVTOT_NH on either side of the pressure solve (ocvtot_nh.f90), on the
// device, for one GPU holding the whole domain (nprocx = nprocy = 1).
//
// All arrays are kept in the caller's Fortran layout with halos (mo_commo1.f90:105-146), i fastest:
//   p-points (IE,JE,K)           a[(i-1) + IE*((j-1) + JE*(k-1))]
//   u-points (I_start:IE,JE,KE)  a[(i-I_start) + NU*((j-1) + JE*(k-1))],   NU = IE-I_start+1
//   v-points (IE,J_start:JE,KE)  a[(i-1) + IE*((j-J_start) + NV*(k-1))],   NV = JE-J_start+1
// One thread per (i, j) column walks k (or per halo line for the exchanges); consecutive threads
// take consecutive i, so every access is coalesced. FP64 throughout, each expression evaluated left
// to right as the Fortran source writes it (no contraction: the library is built -ffp-contract=off).
#include <hip/hip_runtime.h>

#include "nh_ocv.h"

namespace onh {

namespace {

constexpr int OB = 256;  // threads per block

struct Dims {
  int IE, JE, KE, IS, JS, NU, NV;
  __device__ size_t p3(int i, int j, int k) const {
    return (size_t)(i - 1) + (size_t)IE * ((size_t)(j - 1) + (size_t)JE * (size_t)(k - 1));
  }
  __device__ size_t u3(int i, int j, int k) const {
    return (size_t)(i - IS) + (size_t)NU * ((size_t)(j - 1) + (size_t)JE * (size_t)(k - 1));
  }
  __device__ size_t v3(int i, int j, int k) const {
    return (size_t)(i - 1) + (size_t)IE * ((size_t)(j - JS) + (size_t)NV * (size_t)(k - 1));
  }
  __device__ size_t p2(int i, int j) const { return (size_t)(i - 1) + (size_t)IE * (size_t)(j - 1); }
  __device__ size_t u2(int i, int j) const { return (size_t)(i - IS) + (size_t)NU * (size_t)(j - 1); }
  __device__ size_t v2(int i, int j) const { return (size_t)(i - 1) + (size_t)IE * (size_t)(j - JS); }
};

Dims dims_of(const OcvArrays& a) {
  Dims d;
  d.IE = a.IE;
  d.JE = a.JE;
  d.KE = a.KE;
  d.IS = a.i_start;
  d.JS = a.j_start;
  d.NU = a.IE - a.i_start + 1;
  d.NV = a.JE - a.j_start + 1;
  return d;
}

// ocvtot_nh.f90:11-62 on the interior cells i = 2..IE-1, j = 2..JE-1, all k.
__global__ __launch_bounds__(OB) void k_divergence(Dims d, OcvArrays a, double* __restrict__ divg) {
  const int n = blockIdx.x * OB + threadIdx.x;
  const int ni = d.IE - 2, nj = d.JE - 2;
  if (n >= ni * nj) return;
  const int i = 2 + n % ni, j = 2 + n / ni;
  const double dt = a.dt;
  const double dlxp = a.dlxp[d.p2(i, j)], dlyp = a.dlyp[d.p2(i, j)], zo = a.zo[d.p2(i, j)];
  for (int k = 1; k <= d.KE; ++k) {
    const double gu_i = (a.uoo[d.u3(i, j, k)] - a.uoo_1[d.u3(i, j, k)]) / dt;
    const double gu_m = (a.uoo[d.u3(i - 1, j, k)] - a.uoo_1[d.u3(i - 1, j, k)]) / dt;
    const double gv_j = (a.voe[d.v3(i, j, k)] - a.voe_1[d.v3(i, j, k)]) / dt;
    const double gv_m = (a.voe[d.v3(i, j - 1, k)] - a.voe_1[d.v3(i, j - 1, k)]) / dt;
    const double gw_k = (a.wno[d.p3(i, j, k)] - a.wno_1[d.p3(i, j, k)]) / dt;
    const double gw_p = (a.wno[d.p3(i, j, k + 1)] - a.wno_1[d.p3(i, j, k + 1)]) / dt;
    const double divgu = (gu_i - gu_m) / dlxp;
    const double divgv = (gv_m - gv_j) / dlyp;
    const double h = fmax(a.ddpo[d.p3(i, j, k)], 1.0e-3);
    const double divgw = (k == 1) ? (gw_k - gw_p) / (h + zo) : (gw_k - gw_p) / h;
    divg[d.p3(i, j, k)] = divgu + divgv + divgw;
  }
}

// bounds_exch_3d (mo_parallel.f90:551-668), one process, x then y (separate launches):
// x with ICYCLI_X: A(1) = A(IE-1), A(IE) = A(2) for every j, k; y with ICYCLI_Y: A(j0) = A(JE-1),
// A(JE) = A(j0+1), where j0 = 0 for v-point arrays with J_start = 0 and 1 otherwise.
__global__ __launch_bounds__(OB) void k_exch_x(double* __restrict__ a, int IE, int i0, int nj, int nk) {
  const int n = blockIdx.x * OB + threadIdx.x;
  if (n >= nj * nk) return;
  const int ni = IE - i0 + 1;
  double* row = a + (size_t)ni * n;  // row (j, k): A(i) at row[i - i0]
  const double xs1 = row[2 - i0], xs2 = row[IE - 1 - i0];
  row[1 - i0] = xs2;
  row[IE - i0] = xs1;
}

__global__ __launch_bounds__(OB) void k_exch_y(double* __restrict__ a, int JE, int j0, int ni, int nk) {
  const int n = blockIdx.x * OB + threadIdx.x;
  if (n >= ni * nk) return;
  const int i = n % ni, k = n / ni;
  const int nj = JE - j0 + 1;
  double* col = a + i + (size_t)ni * nj * k;  // A(., j, k) at col[(j - j0) * ni]
  const double y1 = col[(size_t)1 * ni], y2 = col[(size_t)(JE - 1 - j0) * ni];
  col[0] = y2;
  col[(size_t)(JE - j0) * ni] = y1;
}

// Open-boundary copies of PNH (ocvtot_nh.f90:154-186): x for j = 2..JE1, then y for every i.
__global__ __launch_bounds__(OB) void k_ob_x(double* __restrict__ p, int IE, int JE, int KE, int west, int east) {
  const int n = blockIdx.x * OB + threadIdx.x;
  const int nj = JE - 2;
  if (n >= nj * KE) return;
  const int j = 2 + n % nj, k = 1 + n / nj;
  double* row = p + (size_t)IE * ((size_t)(j - 1) + (size_t)JE * (size_t)(k - 1));
  if (west) row[0] = row[1];
  if (east) row[IE - 1] = row[IE - 2];
}

__global__ __launch_bounds__(OB) void k_ob_y(double* __restrict__ p, int IE, int JE, int KE, int north, int south) {
  const int n = blockIdx.x * OB + threadIdx.x;
  if (n >= IE * KE) return;
  const int i = 1 + n % IE, k = 1 + n / IE;
  double* col = p + (size_t)(i - 1) + (size_t)IE * JE * (size_t)(k - 1);
  if (north) col[0] = col[(size_t)IE];
  if (south) col[(size_t)IE * (JE - 1)] = col[(size_t)IE * (JE - 2)];
}

// PNH interior (i = 2..IE-1, j = 2..JE-1) from the resident solution (padded interior layout).
__global__ __launch_bounds__(OB) void k_pnh_interior(Dims d, const double* __restrict__ p64, int pitch,
                                                     long long plane, double* __restrict__ pnh) {
  const int n = blockIdx.x * OB + threadIdx.x;
  const int ni = d.IE - 2, nj = d.JE - 2;
  if (n >= ni * nj) return;
  const int ii = n % ni, jj = n / ni;
  for (int k = 1; k <= d.KE; ++k)
    pnh[d.p3(ii + 2, jj + 2, k)] = p64[(size_t)(k - 1) * plane + (size_t)jj * pitch + ii];
}

// ocvtot_nh.f90:189-194: UOO += AMSUO*DTDXUO*(PNH(I)-PNH(I+1)), J = 2..JE1, I = I_start+1..IE1.
__global__ __launch_bounds__(OB) void k_u_correct(Dims d, OcvArrays a, const double* __restrict__ pnh,
                                                  double* __restrict__ uoo) {
  const int n = blockIdx.x * OB + threadIdx.x;
  const int ni = (d.IE - 1) - (d.IS + 1) + 1, nj = d.JE - 2;
  if (n >= ni * nj) return;
  const int i = d.IS + 1 + n % ni, j = 2 + n / ni;
  const double dtdx = a.dtdxuo[d.u2(i, j)];
  for (int k = 1; k <= d.KE; ++k) {
    const size_t u = d.u3(i, j, k);
    uoo[u] = uoo[u] + a.amsuo[u] * dtdx * (pnh[d.p3(i, j, k)] - pnh[d.p3(i + 1, j, k)]);
  }
}

// ocvtot_nh.f90:196-201: VOE += AMSUE*DPYE*(PNH(J+1)-PNH(J)), J = J_start+1..JE1, I = 2..IE1.
__global__ __launch_bounds__(OB) void k_v_correct(Dims d, OcvArrays a, const double* __restrict__ pnh,
                                                  double* __restrict__ voe) {
  const int n = blockIdx.x * OB + threadIdx.x;
  const int ni = d.IE - 2, nj = (d.JE - 1) - (d.JS + 1) + 1;
  if (n >= ni * nj) return;
  const int i = 2 + n % ni, j = d.JS + 1 + n / ni;
  const double dpy = a.dpye[d.v2(i, j)];
  for (int k = 1; k <= d.KE; ++k) {
    const size_t v = d.v3(i, j, k);
    voe[v] = voe[v] + a.amsue[v] * dpy * (pnh[d.p3(i, j + 1, k)] - pnh[d.p3(i, j, k)]);
  }
}

// ocvtot_nh.f90:207-228: WNO = 0, then for K = KE..1, J = J_start+1..JE, I = I_start+1..IE
// WNO(K) = WNO(K+1) + DTI*WETO*(DTDXPO*(UOO(I-1)*DDUO(I-1)*DLYU(I-1) - UOO(I)*DDUO(I)*DLYU(I))/DLYP)
//                   + DTI*WETO*(+DTDYO*(VOE(J)*DDUE(J)*DLXV(J) - VOE(J-1)*DDUE(J-1)*DLXV(J-1))/DLXP).
// The caller zeroes wno first; the columns outside the loop range stay 0.
__global__ __launch_bounds__(OB) void k_wno(Dims d, OcvArrays a, const double* __restrict__ uoo,
                                            const double* __restrict__ voe, double* __restrict__ wno) {
  const int n = blockIdx.x * OB + threadIdx.x;
  const int ni = d.IE - d.IS, nj = d.JE - d.JS;
  if (n >= ni * nj) return;
  const int i = d.IS + 1 + n % ni, j = d.JS + 1 + n / ni;
  const double dtdxpo = a.dtdxpo[d.p2(i, j)], dtdyo = a.dtdyo[d.p2(i, j)];
  const double dlxp = a.dlxp[d.p2(i, j)], dlyp = a.dlyp[d.p2(i, j)];
  const double dlyu_m = a.dlyu[d.u2(i - 1, j)], dlyu_i = a.dlyu[d.u2(i, j)];
  const double dlxv_j = a.dlxv[d.v2(i, j)], dlxv_m = a.dlxv[d.v2(i, j - 1)];
  double w = 0.0;  // WNO(KEP) = 0
  for (int k = d.KE; k >= 1; --k) {
    const double dw = a.dti * a.weto[d.p3(i, j, k)];
    const double fx =
        dtdxpo * (uoo[d.u3(i - 1, j, k)] * a.dduo[d.u3(i - 1, j, k)] * dlyu_m - uoo[d.u3(i, j, k)] * a.dduo[d.u3(i, j, k)] * dlyu_i) /
        dlyp;
    const double fy =
        dtdyo * (voe[d.v3(i, j, k)] * a.ddue[d.v3(i, j, k)] * dlxv_j - voe[d.v3(i, j - 1, k)] * a.ddue[d.v3(i, j - 1, k)] * dlxv_m) /
        dlxp;
    w = w + dw * fx + dw * fy;
    wno[d.p3(i, j, k)] = w;
  }
}

inline unsigned blocks(long long n) { return (unsigned)((n + OB - 1) / OB); }

}  // namespace

hipError_t launch_divergence(hipStream_t st, const OcvArrays& a, double* divg) {
  const Dims d = dims_of(a);
  const long long n = (long long)(a.IE - 2) * (a.JE - 2);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_divergence, dim3(blocks(n)), dim3(OB), 0, st, d, a, divg);
  return hipGetLastError();
}

hipError_t launch_exchange(hipStream_t st, double* arr, int IE, int JE, int nk, int i0, int j0, int cyc_x, int cyc_y) {
  const int ni = IE - i0 + 1, nj = JE - j0 + 1;
  if (cyc_x) {
    hipLaunchKernelGGL(k_exch_x, dim3(blocks((long long)nj * nk)), dim3(OB), 0, st, arr, IE, i0, nj, nk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (cyc_y) {
    hipLaunchKernelGGL(k_exch_y, dim3(blocks((long long)ni * nk)), dim3(OB), 0, st, arr, JE, j0, ni, nk);
    return hipGetLastError();
  }
  return hipSuccess;
}

hipError_t launch_project(hipStream_t st, const OcvArrays& a, const double* p64, int pitch, long long plane,
                          double* pnh, double* uoo, double* voe, double* wno) {
  const Dims d = dims_of(a);
  hipError_t e = hipSuccess;
#define OCV_CHK(x)              \
  do {                          \
    e = (x);                    \
    if (e != hipSuccess) return e; \
  } while (0)
  const long long nint = (long long)(a.IE - 2) * (a.JE - 2);
  hipLaunchKernelGGL(k_pnh_interior, dim3(blocks(nint)), dim3(OB), 0, st, d, p64, pitch, plane, pnh);
  OCV_CHK(hipGetLastError());
  // bounds_exch(PNH) (:151), then the open-boundary copies (:154-186)
  OCV_CHK(launch_exchange(st, pnh, a.IE, a.JE, a.KE, 1, 1, a.cyc_x, a.cyc_y));
  if (!a.cyc_x && (a.open_w || a.open_e)) {
    hipLaunchKernelGGL(k_ob_x, dim3(blocks((long long)(a.JE - 2) * a.KE)), dim3(OB), 0, st, pnh, a.IE, a.JE, a.KE,
                       a.open_w, a.open_e);
    OCV_CHK(hipGetLastError());
  }
  if (!a.cyc_y && (a.open_n || a.open_s)) {
    hipLaunchKernelGGL(k_ob_y, dim3(blocks((long long)a.IE * a.KE)), dim3(OB), 0, st, pnh, a.IE, a.JE, a.KE,
                       a.open_n, a.open_s);
    OCV_CHK(hipGetLastError());
  }
  const long long nu = (long long)((a.IE - 1) - (a.i_start + 1) + 1) * (a.JE - 2);
  hipLaunchKernelGGL(k_u_correct, dim3(blocks(nu)), dim3(OB), 0, st, d, a, pnh, uoo);
  OCV_CHK(hipGetLastError());
  OCV_CHK(launch_exchange(st, uoo, a.IE, a.JE, a.KE, a.i_start, 1, a.cyc_x, a.cyc_y));
  const long long nv = (long long)(a.IE - 2) * ((a.JE - 1) - (a.j_start + 1) + 1);
  hipLaunchKernelGGL(k_v_correct, dim3(blocks(nv)), dim3(OB), 0, st, d, a, pnh, voe);
  OCV_CHK(hipGetLastError());
  OCV_CHK(launch_exchange(st, voe, a.IE, a.JE, a.KE, 1, a.j_start, a.cyc_x, a.cyc_y));
  OCV_CHK(hipMemsetAsync(wno, 0, sizeof(double) * (size_t)a.IE * a.JE * (a.KE + 1), st));
  const long long nw = (long long)(a.IE - a.i_start) * (a.JE - a.j_start);
  hipLaunchKernelGGL(k_wno, dim3(blocks(nw)), dim3(OB), 0, st, d, a, uoo, voe, wno);
  OCV_CHK(hipGetLastError());
  OCV_CHK(launch_exchange(st, wno, a.IE, a.JE, a.KE + 1, 1, 1, a.cyc_x, a.cyc_y));
#undef OCV_CHK
  return hipSuccess;
}

}  // namespace onh
