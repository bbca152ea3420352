// nh_sor.hip -- ORCTM's barotropic free-surface solve on the device: the coefficients of ITPREP
// (itprep.f90:15-38) and the right-hand side and red-black SOR iteration of TRONEU
// (troneu.f90:48-131), for one GPU holding the whole domain (p_ioff = p_joff = 0).
//
// Arrays are kept in the caller's Fortran layout with halos (p-points (IE,JE), u-points
// (I_start:IE,JE), v-points (IE,J_start:JE)). The first half-sweep reads the halo as TRONEU initialises
// it (B1O/XX), so it runs on its own (k_sor, one thread per cell) and is followed by the exchange of the
// cyclic halo (bounds_exch). From then on the halo holds copies of interior cells, and the remaining
// half-sweeps run temporally blocked (nh_sor_tb.h: several half-sweeps per launch, staged in LDS, the
// cyclic halo read through a periodic index map), ending with one exchange. Within a colour every
// update reads only values from before the half-sweep, so every schedule gives the reference's bits.
// The whole sequence is captured once into a hipGraph and replayed per call; the half-sweeps per launch
// are chosen by timing the compiled depths at creation. FP64 throughout, expressions left to right as
// the Fortran source writes them (no contraction).
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "../../include/orctm_nh.h"
#include "nh_lock.h"
#include "nh_ocv.h"
#include "nh_sor_tb.h"

namespace onh {
namespace {

constexpr int SB = 256;

struct SorDims {
  int IE, JE, IS, JS, NU;
  __device__ size_t p2(int i, int j) const { return (size_t)(i - 1) + (size_t)IE * (size_t)(j - 1); }
  __device__ size_t u2(int i, int j) const { return (size_t)(i - IS) + (size_t)NU * (size_t)(j - 1); }
  __device__ size_t v2(int i, int j) const { return (size_t)(i - 1) + (size_t)IE * (size_t)(j - JS); }
};

struct SorArrays {
  const double *dlxp, *dlyp, *dtdxpo, *dtdyo, *weto1, *dlxu, *dlyu, *deuto, *dlxv, *dlyv, *deute;
  double *xx, *uf, *vf, *ff, *b1o, *z1o, *zb, *u1o, *uzo, *v1e, *vze;  // zb: second sweep buffer
};

inline unsigned nblk(long long n) { return (unsigned)((n + SB - 1) / SB); }

// itprep.f90:15-38
__global__ __launch_bounds__(SB) void k_itprep_p(SorDims d, SorArrays a) {
  const int n = blockIdx.x * SB + threadIdx.x;
  if (n >= d.IE * d.JE) return;
  const int i = 1 + n % d.IE, j = 1 + n / d.IE;
  a.xx[d.p2(i, j)] = a.dlxp[d.p2(i, j)] * a.dlyp[d.p2(i, j)];
  a.ff[d.p2(i, j)] = 0.0;
}

__global__ __launch_bounds__(SB) void k_itprep_uv(SorDims d, SorArrays a, double g, double conn, double stabn,
                                                  double dt2) {
  const int n = blockIdx.x * SB + threadIdx.x;
  const int nu = d.NU * d.JE, nv = d.IE * (d.JE - d.JS + 1);
  if (n < nu) {
    const int i = d.IS + n % d.NU, j = 1 + n / d.NU;
    const size_t u = d.u2(i, j);
    a.uf[u] = a.deuto[u] * a.dlyu[u] / a.dlxu[u] * g * conn * stabn * dt2;
  } else if (n < nu + nv) {
    const int m = n - nu;
    const int i = 1 + m % d.IE, j = d.JS + m / d.IE;
    const size_t v = d.v2(i, j);
    a.vf[v] = a.deute[v] * a.dlxv[v] / a.dlyv[v] * g * conn * stabn * dt2;
  }
}

__global__ __launch_bounds__(SB) void k_itprep_ff(SorDims d, SorArrays a) {
  const int ni = d.IE - d.IS, nj = d.JE - d.JS;
  const int n = blockIdx.x * SB + threadIdx.x;
  if (n >= ni * nj) return;
  const int i = d.IS + 1 + n % ni, j = d.JS + 1 + n / ni;
  a.ff[d.p2(i, j)] =
      1. / (a.xx[d.p2(i, j)] + a.uf[d.u2(i, j)] + a.uf[d.u2(i - 1, j)] + a.vf[d.v2(i, j)] + a.vf[d.v2(i, j - 1)]);
}

// troneu.f90:48-65: B1O for J = J_start+1..JE, I = I_start+1..IE.
__global__ __launch_bounds__(SB) void k_b1o(SorDims d, SorArrays a, double cono, double conn) {
  const int ni = d.IE - d.IS, nj = d.JE - d.JS;
  const int n = blockIdx.x * SB + threadIdx.x;
  if (n >= ni * nj) return;
  const int i = d.IS + 1 + n % ni, j = d.JS + 1 + n / ni;
  const double fx = a.dtdxpo[d.p2(i, j)] *
                    ((cono * a.u1o[d.u2(i - 1, j)] + a.uzo[d.u2(i - 1, j)] * conn) * a.dlyu[d.u2(i - 1, j)] -
                     (a.u1o[d.u2(i, j)] * cono + a.uzo[d.u2(i, j)] * conn) * a.dlyu[d.u2(i, j)]) /
                    a.dlyp[d.p2(i, j)];
  const double fy = a.dtdyo[d.p2(i, j)] *
                    ((a.v1e[d.v2(i, j)] * cono + conn * a.vze[d.v2(i, j)]) * a.dlxv[d.v2(i, j)] -
                     (a.v1e[d.v2(i, j - 1)] * cono + conn * a.vze[d.v2(i, j - 1)]) * a.dlxv[d.v2(i, j - 1)]) /
                    a.dlxp[d.p2(i, j)];
  const double b = a.weto1[d.p2(i, j)] * (fx + fy);
  a.b1o[d.p2(i, j)] = b * a.xx[d.p2(i, j)];
}

// troneu.f90:67-101: Z1O = 0, boundary rows and columns B1O/XX (in both sweep buffers).
__global__ __launch_bounds__(SB) void k_z1o_init(SorDims d, SorArrays a) {
  const int n = blockIdx.x * SB + threadIdx.x;
  if (n >= d.IE * d.JE) return;
  const int i = 1 + n % d.IE, j = 1 + n / d.IE;
  const size_t p = d.p2(i, j);
  const double z = (i == 1 || i == d.IE || j == 1 || j == d.JE) ? a.b1o[p] / a.xx[p] : 0.0;
  a.z1o[p] = z;
  a.zb[p] = z;
}

// troneu.f90:105-127: one half-sweep over the cells with (i + j + colour) even, i = 2..IE-1,
// j = 2..JE-1 (the reference's IS = MOD(j + colour, 2) + 2 stepping by 2).
__global__ __launch_bounds__(SB) void k_sor(SorDims d, SorArrays a, double* z, int colour, double sorpar,
                                            double sorpai) {
  const int ni = d.IE - 2, nj = d.JE - 2;
  const int n = blockIdx.x * SB + threadIdx.x;
  if (n >= ni * nj) return;
  const int i = 2 + n % ni, j = 2 + n / ni;
  if ((i + j + colour) & 1) return;
  const size_t p = d.p2(i, j);
  const double s = a.b1o[p] + a.uf[d.u2(i, j)] * z[d.p2(i + 1, j)] + a.uf[d.u2(i - 1, j)] * z[d.p2(i - 1, j)] +
                   a.vf[d.v2(i, j)] * z[d.p2(i, j + 1)] + a.vf[d.v2(i, j - 1)] * z[d.p2(i, j - 1)];
  z[p] = sorpar * a.ff[p] * s + sorpai * z[p];
}

// nh half-sweeps, colours c0, c0+1, ..., on the tile of this block (nh_sor_tb.h); zin and zout are the
// two sweep buffers.
template <int T, int RX, int RY, int NT, bool SEAM>
__global__ __launch_bounds__(NT) void k_sor_tb(TbGeom g, const double* __restrict__ zin, double* __restrict__ zout,
                                               const double* __restrict__ b1o, const double* __restrict__ ff,
                                               const double* __restrict__ uf, const double* __restrict__ vf, int nh,
                                               int c0, double sorpar, double sorpai) {
  using K = SorTb<T, RX, RY, NT, SEAM>;
  __shared__ double lz[K::LDS];
  int vx0, vy0;
  K::origin(g, tb_tile(blockIdx.x, gridDim.x), vx0, vy0);
  typename K::Regs r;
  const int tid = threadIdx.x;
  K::load(tid, vx0, vy0, g, zin, b1o, ff, uf, vf, sorpar, lz, r);
  __syncthreads();
#define ONH_HALF(C)                                  \
  do {                                               \
    K::template compute<C>(tid, sorpai, lz, r);      \
    __syncthreads();                                 \
    K::template write<C>(tid, lz, r);                \
    __syncthreads();                                 \
  } while (0)
  int h = 0;
  if ((c0 & 1) && nh > 0) {
    ONH_HALF(1);
    h = 1;
  }
  for (; h + 2 <= nh; h += 2) {
    ONH_HALF(0);
    ONH_HALF(1);
  }
  if (h < nh) ONH_HALF(0);
#undef ONH_HALF
  K::store(tid, vx0, vy0, g, lz, zout);
}

constexpr int kTbDepths[] = {
#define X(T, RX, RY, NT) T,
    ORCTM_SOR_TB_VARIANTS(X)
#undef X
};

}  // namespace
}  // namespace onh

using namespace onh;

struct orctm_nh_sor {
  orctm_nh_sor_grid_t g{};
  SorDims d{};
  SorArrays a{};
  std::vector<void*> allocs;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int depth = 0;     // half-sweeps per blocked launch
  int launches = 0;  // kernel launches of the sweeps (the first half-sweep included)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex mu;
};

namespace {

void sor_free(orctm_nh_sor_t* s) {
  if (s->graph) hipGraphExecDestroy(s->graph);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->stream) hipStreamDestroy(s->stream);
  for (void* p : s->allocs) hipFree(p);
  s->allocs.clear();
}

int sor_map(hipError_t e) {
  if (e == hipSuccess) return ORCTM_NH_OK;
  if (e == hipErrorOutOfMemory) return ORCTM_NH_ERR_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice || e == hipErrorNoBinaryForGpu) return ORCTM_NH_ERR_NODEVICE;
  return ORCTM_NH_ERR_HIP;
}

// The sweeps (iterations >= 1): the first half-sweep on its own, then 2*iterations - 1 half-sweeps
// from colour 1, T per blocked launch, alternating between the two buffers so that the last launch
// writes z1o, then the final exchange.
template <int T, int RX, int RY, int NT, bool SEAM>
hipError_t sor_blocked(orctm_nh_sor_t* s, int& launches) {
  using K = SorTb<T, RX, RY, NT, SEAM>;
  const SorDims& d = s->d;
  TbGeom g{d.IE, d.JE, d.IS, d.JS, d.NU, s->g.cyclic_x, s->g.cyclic_y, 0};
  g.tiles_x = (d.IE - 2 + K::TXo - 1) / K::TXo;
  const long long tiles_y = (d.JE - 2 + K::TYo - 1) / K::TYo, nb = (long long)g.tiles_x * tiles_y;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  const long long H = 2LL * s->g.iterations - 1, nl = (H + T - 1) / T;
  double* buf[2] = {s->a.z1o, s->a.zb};
  const int s0 = (int)(nl & 1);
  const double sorpar = s->g.sorpar, sorpai = 1. - s->g.sorpar;
  const long long n = (long long)(d.IE - 2) * (d.JE - 2);
  hipLaunchKernelGGL(k_sor, dim3(nblk(n)), dim3(SB), 0, s->stream, d, s->a, buf[s0], 0, sorpar, sorpai);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = launch_exchange(s->stream, buf[s0], d.IE, d.JE, 1, 1, 1, s->g.cyclic_x, s->g.cyclic_y);
  launches = 1;
  for (long long l = 0; l < nl && e == hipSuccess; ++l) {
    const int nh = (int)(H - l * T < T ? H - l * T : T);
    const int c0 = (int)((1 + l * T) & 1);
    hipLaunchKernelGGL((k_sor_tb<T, RX, RY, NT, SEAM>), dim3((unsigned)nb), dim3(NT), 0, s->stream, g,
                       (const double*)buf[(s0 + l) & 1], buf[(s0 + l + 1) & 1], (const double*)s->a.b1o,
                       (const double*)s->a.ff, (const double*)s->a.uf, (const double*)s->a.vf, nh, c0, sorpar, sorpai);
    e = hipGetLastError();
    ++launches;
  }
  if (e == hipSuccess) e = launch_exchange(s->stream, s->a.z1o, d.IE, d.JE, 1, 1, 1, s->g.cyclic_x, s->g.cyclic_y);
  return e;
}

hipError_t sor_sweeps(orctm_nh_sor_t* s, int T, int& launches) {
  // an odd period in a cyclic direction puts cells of one colour side by side across the seam
  const bool seam = (s->g.cyclic_x && ((s->d.IE - 2) & 1)) || (s->g.cyclic_y && ((s->d.JE - 2) & 1));
#define X(TT, RX, RY, NT) \
  if (T == TT) return seam ? sor_blocked<TT, RX, RY, NT, true>(s, launches) : sor_blocked<TT, RX, RY, NT, false>(s, launches);
  ORCTM_SOR_TB_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t sor_capture(orctm_nh_sor_t* s, int T, hipGraphExec_t* out, int& launches) {
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t e1 = sor_sweeps(s, T, launches);
  const hipError_t e2 = hipStreamEndCapture(s->stream, &graph);
  e = e1 != hipSuccess ? e1 : e2;
  if (e == hipSuccess) e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  return e;
}

// Captures the sweeps for the requested depth, or times every compiled depth on the resident buffers
// (the values do not change the work) and keeps the fastest. The caller holds the device lock.
hipError_t sor_prepare(orctm_nh_sor_t* s, int want) {
  if (s->g.iterations < 1) return hipSuccess;
  hipError_t e = hipSuccess;
  float best = 0.f;
  for (int T : kTbDepths) {
    if (want && T != want) continue;
    hipGraphExec_t ge = nullptr;
    int launches = 0;
    e = sor_capture(s, T, &ge, launches);
    float ms = 0.f;
    if (e == hipSuccess && !want) {
      e = hipGraphLaunch(ge, s->stream);  // warm
      if (e == hipSuccess) e = hipEventRecord(s->ev0, s->stream);
      for (int r = 0; r < 3 && e == hipSuccess; ++r) e = hipGraphLaunch(ge, s->stream);
      if (e == hipSuccess) e = hipEventRecord(s->ev1, s->stream);
      if (e == hipSuccess) e = hipEventSynchronize(s->ev1);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->ev0, s->ev1);
    }
    if (e != hipSuccess) {
      if (ge) hipGraphExecDestroy(ge);
      return e;
    }
    if (!s->graph || ms < best) {
      if (s->graph) hipGraphExecDestroy(s->graph);
      s->graph = ge;
      s->depth = T;
      s->launches = launches;
      best = ms;
    } else {
      hipGraphExecDestroy(ge);
    }
  }
  return s->graph ? hipSuccess : hipErrorInvalidValue;
}

}  // namespace

extern "C" {

int orctm_nh_sor_create(const orctm_nh_sor_grid_t* g, orctm_nh_sor_t** out) {
  if (!out || !g) return ORCTM_NH_ERR_INVALID;
  *out = nullptr;
  const double* need[] = {g->dlxp, g->dlyp, g->dtdxpo, g->dtdyo, g->weto, g->dlxu,
                          g->dlyu, g->deuto, g->dlxv, g->dlyv, g->deute};
  for (const double* p : need)
    if (!p) return ORCTM_NH_ERR_INVALID;
  if (g->ie < 3 || g->je < 3 || g->i_start < 0 || g->i_start > 1 || g->j_start < 0 || g->j_start > 1 ||
      g->iterations < 0 || !(g->sorpar > 0.0 && g->sorpar < 2.0) || (long long)g->ie * g->je > (1LL << 30))
    return ORCTM_NH_ERR_INVALID;
  bool depth_ok = g->half_sweeps == 0;
  for (int T : kTbDepths) depth_ok = depth_ok || g->half_sweeps == T;
  if (!depth_ok) return ORCTM_NH_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || g->device < 0 || g->device >= ndev) return ORCTM_NH_ERR_NODEVICE;
  std::lock_guard<std::recursive_mutex> dev(device_lock());
  if (hipSetDevice(g->device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  orctm_nh_sor_t* s = new (std::nothrow) orctm_nh_sor_t();
  if (!s) return ORCTM_NH_ERR_NOMEM;
  s->g = *g;
  SorDims& d = s->d;
  d.IE = g->ie;
  d.JE = g->je;
  d.IS = g->i_start;
  d.JS = g->j_start;
  d.NU = g->ie - g->i_start + 1;
  const size_t p2 = (size_t)g->ie * g->je, u2 = (size_t)d.NU * g->je, v2 = (size_t)g->ie * (g->je - g->j_start + 1);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&s->ev0);
  if (e == hipSuccess) e = hipEventCreate(&s->ev1);
  auto alloc = [&](double*& dst, size_t n, const double* src) {
    void* p = nullptr;
    if (e == hipSuccess) e = hipMalloc(&p, n * sizeof(double));
    if (e == hipSuccess) s->allocs.push_back(p);
    if (e == hipSuccess) e = src ? hipMemcpy(p, src, n * sizeof(double), hipMemcpyHostToDevice)
                                 : hipMemset(p, 0, n * sizeof(double));
    dst = static_cast<double*>(p);
  };
  auto cput = [&](const double*& dst, size_t n, const double* src) {
    double* q = nullptr;
    alloc(q, n, src);
    dst = q;
  };
  SorArrays& a = s->a;
  cput(a.dlxp, p2, g->dlxp);
  cput(a.dlyp, p2, g->dlyp);
  cput(a.dtdxpo, p2, g->dtdxpo);
  cput(a.dtdyo, p2, g->dtdyo);
  cput(a.weto1, p2, g->weto);  // WETO(:,:,1): the first IE*JE values of WETO(IE,JE,KE)
  cput(a.dlxu, u2, g->dlxu);
  cput(a.dlyu, u2, g->dlyu);
  cput(a.deuto, u2, g->deuto);
  cput(a.dlxv, v2, g->dlxv);
  cput(a.dlyv, v2, g->dlyv);
  cput(a.deute, v2, g->deute);
  alloc(a.xx, p2, nullptr);
  alloc(a.uf, u2, nullptr);
  alloc(a.vf, v2, nullptr);
  alloc(a.ff, p2, nullptr);
  alloc(a.b1o, p2, nullptr);
  alloc(a.z1o, p2, nullptr);
  alloc(a.zb, p2, nullptr);
  alloc(a.u1o, u2, nullptr);
  alloc(a.uzo, u2, nullptr);
  alloc(a.v1e, v2, nullptr);
  alloc(a.vze, v2, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();  // the null-stream copies and fills above
  if (e == hipSuccess) {
    const double dt2 = g->dt * g->dt;
    hipLaunchKernelGGL(k_itprep_p, dim3(nblk((long long)p2)), dim3(SB), 0, s->stream, d, a);
    e = hipGetLastError();
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_itprep_uv, dim3(nblk((long long)(u2 + v2))), dim3(SB), 0, s->stream, d, a, g->g, g->conn,
                         g->stabn, dt2);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_itprep_ff, dim3(nblk((long long)(d.IE - d.IS) * (d.JE - d.JS))), dim3(SB), 0, s->stream,
                         d, a);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = launch_exchange(s->stream, a.ff, d.IE, d.JE, 1, 1, 1, g->cyclic_x, g->cyclic_y);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  }
  if (e == hipSuccess) e = sor_prepare(s, g->half_sweeps);
  if (e != hipSuccess) {
    sor_free(s);
    delete s;
    return sor_map(e);
  }
  *out = s;
  return ORCTM_NH_OK;
}

int orctm_nh_troneu(orctm_nh_sor_t* s, const double* u1o, const double* uzo, const double* v1e, const double* vze,
                    double* b1o, double* z1o, double* device_ms) {
  if (!s || !u1o || !uzo || !v1e || !vze || !b1o || !z1o) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  hipSetDevice(s->g.device);
  const SorDims& d = s->d;
  const size_t p2 = (size_t)d.IE * d.JE, u2 = (size_t)d.NU * d.JE, v2 = (size_t)d.IE * (d.JE - d.JS + 1);
  hipStream_t st = s->stream;
  hipError_t e = hipMemcpyAsync(s->a.u1o, u1o, u2 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->a.uzo, uzo, u2 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->a.v1e, v1e, v2 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->a.vze, vze, v2 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->a.b1o, b1o, p2 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipEventRecord(s->ev0, st);
  if (e == hipSuccess) {
    const long long nb = (long long)(d.IE - d.IS) * (d.JE - d.JS);
    hipLaunchKernelGGL(k_b1o, dim3(nblk(nb)), dim3(SB), 0, st, d, s->a, s->g.cono, s->g.conn);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_z1o_init, dim3(nblk((long long)p2)), dim3(SB), 0, st, d, s->a);
    e = hipGetLastError();
  }
  if (e == hipSuccess && s->graph) e = hipGraphLaunch(s->graph, st);
  if (e == hipSuccess) e = hipEventRecord(s->ev1, st);
  if (e == hipSuccess) e = hipMemcpyAsync(z1o, s->a.z1o, p2 * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(b1o, s->a.b1o, p2 * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && device_ms) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, s->ev0, s->ev1);
    *device_ms = ms;
  }
  return sor_map(e);
}

int orctm_nh_sor_info(orctm_nh_sor_t* s, int32_t* half_sweeps, int32_t* launches) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  if (half_sweeps) *half_sweeps = s->depth;
  if (launches) *launches = s->launches;
  return ORCTM_NH_OK;
}

int orctm_nh_sor_destroy(orctm_nh_sor_t* s) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::recursive_mutex> dev(device_lock());
  hipSetDevice(s->g.device);
  if (s->stream) hipStreamSynchronize(s->stream);
  sor_free(s);
  delete s;
  return ORCTM_NH_OK;
}

}  // extern "C"
