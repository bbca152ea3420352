// nh_api.hip -- C ABI of the gfx950 nonhydrostatic pressure Poisson solver (include/orctm_nh.h).
//
// A handle owns one system's device data: the assembled reference operator (FP64), the scaled
// inner operator and line-preconditioner factors (FP32 or FP64), the right-hand side, the FP64
// solution kept resident for warm starts (KSPSetInitialGuessNonzero, mo_petsc.f90:232) and the
// inner work vectors. A solver context (Ctx) drives one or more systems through the same kernel
// launches: FP64 residual -> inner PCG in fixed-length chunks captured in a hipGraph (convergence
// polled between chunks) -> FP64 refinement update, until every system has converged.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/orctm_nh.h"
#include "nh_bot.h"
#include "nh_kernels.h"
#include "nh_lock.h"
#include "nh_tiles.h"
#include "nh_ocv.h"
#include "nh_internal.h"
#include "nh_pool.h"

using namespace onh;

std::recursive_mutex& onh::device_lock() {
  static std::recursive_mutex mu;
  return mu;
}

namespace {

// Tiles of the inner-iteration kernels for one tile height (systems grouped by their tuned height).
struct Group {
  int h = 12;
  int ntiles = 0;
  Tile* d_tiles = nullptr;
  int npairs = 0;
  Tile* d_pairs = nullptr;  // the tiles with an even bx, for the column update over tile pairs
};

struct Ctx {
  int nsys = 0;
  bool fp32 = true;
  int ntiles = 0;  // set-up tiling (64 x TY), d_tiles
  int max_outer = 0;
  int maxit = 10000;
  int check_every = 16;
  int device = 0;
  Sys* d_sys = nullptr;
  Tile* d_tiles = nullptr;
  int n_otiles = 0;             // tiling of the refinement kernels (64 x kOuterRows), d_otiles
  Tile* d_otiles = nullptr;
  Outer* d_outer = nullptr;
  void* d_inner = nullptr;
  Outer* h_outer = nullptr;  // pinned
  void* h_inner = nullptr;   // pinned
  unsigned long long* d_seeds = nullptr;  // per system: seed of the right-hand-side perturbation
  unsigned long long* h_seeds = nullptr;  // pinned staging of d_seeds
  hipStream_t stream = nullptr;
  bool own_stream = true;          // false: the stream belongs to the handle the context solves
  bool use_graphs = true;          // false: chunks are launched directly (short-lived serving cohorts)
  std::vector<Group> groups;       // inner-iteration tilings, one per tile height in use
  std::vector<int> heights;        // tile height of each system
  std::vector<int> tab;            // per system: 1 = ZW/ZP read from the table and window
  bool buf = false;                // the matvec kernel addresses memory through buffer descriptors
  std::map<int, hipGraphExec_t> graphs;  // captured runs of 1, 2, 4, ... inner iterations
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // optional kernel timing: events around every matvec / update launch (direct launches)
  bool prof = false;
  std::vector<hipEvent_t> kev;      // 4 per iteration of a chunk
  std::vector<int> iter_seen;       // per system, inner iterations already accounted
  std::vector<double> lrate;        // per system, mean log-reduction of r.r per inner iteration (0: unknown)
  double mv_ms = 0.0, up_ms = 0.0;
  long long mv_n = 0, up_n = 0, cells = 0;
};

size_t inner_size(bool fp32) { return fp32 ? sizeof(Inner<float>) : sizeof(Inner<double>); }

int inner_iter(const Ctx& c, int q) {
  if (c.fp32) return reinterpret_cast<const Inner<float>*>(c.h_inner)[q].iter;
  return reinterpret_cast<const Inner<double>*>(c.h_inner)[q].iter;
}

int inner_done(const Ctx& c, int q) {
  if (c.fp32) return reinterpret_cast<const Inner<float>*>(c.h_inner)[q].done;
  return reinterpret_cast<const Inner<double>*>(c.h_inner)[q].done;
}

}  // namespace

// Device arrays of one assembled operator (compact form: ZW, ZP, flag bits, per-column metric
// products, their inner-precision copies, line-preconditioner factors, max S). Shared, read only,
// by every handle cloned from the one that assembled it.
struct OpData {
  int device = 0;
  std::vector<void*> allocs;
  size_t bytes = 0;
  ~OpData() {
    for (void* p : allocs) pool_free(device, p);
  }
};

struct orctm_nh {
  orctm_nh_options_t opts;
  std::shared_ptr<OpData> op;
  int IE = 0, JE = 0, KE = 0, nx = 0, ny = 0, nz = 0, pitch = 0;
  long long plane = 0;
  bool fp32 = true;
  bool guess_set = false;
  Sys sys{};  // host copy, pointers into device memory
  hipStream_t stream = nullptr;    // the handle's own stream: its solver context and every copy use it
  bool ctx_ready = false;          // the single-system context is built on first use (clones skip it)
  std::vector<void*> allocs;
  size_t bytes = 0;
  double* d_fort = nullptr;        // staging for a Fortran (IE,JE,KE) array
  double* d_nb = nullptr;          // ||b|| written by k_rhs
  unsigned long long* d_smax = nullptr;
  Ctx ctx;                         // single-system context
  // grid flags kept for the steps around the solve (orctm_nh_divergence / orctm_nh_project)
  int i_start = 1, j_start = 1, cyc_x = 0, cyc_y = 0, open_w = 0, open_e = 0, open_n = 0, open_s = 0;
  bool ocv_ready = false;
  OcvArrays ocv{};                 // metrics attached by orctm_nh_attach_vgrid + velocity buffers
  double *d_uoo = nullptr, *d_uoo_1 = nullptr, *d_voe = nullptr, *d_voe_1 = nullptr;
  double *d_wno = nullptr, *d_wno_1 = nullptr, *d_zo = nullptr;
  std::unique_ptr<BotState> bot;   // the 2-D bottom-pressure solve (options.bottom_solve)
  std::mutex mu;
};

struct orctm_nh_batch {
  std::vector<orctm_nh_t*> members;
  Ctx ctx;
};

namespace {

thread_local char g_last_error[256] = "";

// Graph capture on one thread and a null-stream, allocation or device-wide call on another do not
// mix: the runtime fails the other thread's call with hipErrorStreamCaptureImplicit even in
// thread-local capture mode. Every capture holds this lock, and so does every entry point that
// allocates, frees, copies on the null stream or synchronises the device, so that handles can be
// created and used from several threads (the serving loop runs its own thread).
#define DEVLOCK std::lock_guard<std::recursive_mutex> dev_lock_(onh::device_lock())

hipError_t note_error(hipError_t e, const char* what, int line) {
  if (e != hipSuccess)
    snprintf(g_last_error, sizeof(g_last_error), "%s (%d) at nh_api.hip:%d: %s", hipGetErrorName(e), (int)e, line,
             what);
  return e;
}

#define HIPCHK(x)                                                   \
  do {                                                              \
    hipError_t e_ = (x);                                            \
    if (e_ != hipSuccess) return note_error(e_, #x, __LINE__);      \
  } while (0)

int map_err(hipError_t e) {
  if (e == hipSuccess) return ORCTM_NH_OK;
  if (g_last_error[0] == 0) note_error(e, "device call", 0);
  if (e == hipErrorOutOfMemory) return ORCTM_NH_ERR_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice || e == hipErrorNoBinaryForGpu) return ORCTM_NH_ERR_NODEVICE;
  return ORCTM_NH_ERR_HIP;
}

// Device memory of a handle comes from the library's pool and is zero-filled on the handle's own
// stream, which every later use of it follows (no device-wide synchronisation, so handles can be
// created and destroyed while other contexts run).
hipError_t dalloc(orctm_nh_t* h, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = pool_alloc(h->opts.device, bytes, p);
  if (e != hipSuccess) return e;
  h->allocs.push_back(*p);
  h->bytes += bytes;
  e = hipMemsetAsync(*p, 0, bytes, h->stream);
  return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}

hipError_t opalloc(orctm_nh_t* h, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = pool_alloc(h->opts.device, bytes, p);
  if (e != hipSuccess) return e;
  h->op->allocs.push_back(*p);
  h->op->bytes += bytes;
  e = hipMemsetAsync(*p, 0, bytes, h->stream);
  return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}

// Per-handle state: right-hand side, solution, inner work vectors, block partials, ||b||.
int alloc_state(orctm_nh_t* h) {
  const size_t NP = (size_t)h->plane * h->nz;
  const size_t es = h->fp32 ? sizeof(float) : sizeof(double);
  Sys& s = h->sys;
  void* p;
  hipError_t e;
#define SALLOC(field, bytes, type)             \
  do {                                         \
    e = dalloc(h, &p, (bytes));                \
    if (e != hipSuccess) return map_err(e);    \
    s.field = reinterpret_cast<type>(p);       \
  } while (0)
  SALLOC(b, NP * sizeof(double), double*);
  SALLOC(p64, NP * sizeof(double), double*);
  SALLOC(x, NP * es, void*);
  SALLOC(r, NP * es, void*);
  SALLOC(z, NP * es, void*);
  SALLOC(pa, NP * es, void*);
  SALLOC(pb, NP * es, void*);
  SALLOC(q, NP * es, void*);
  SALLOC(partials, 2 * (size_t)s.nslots * sizeof(double), double*);
  SALLOC(nbp, sizeof(double), double*);
#undef SALLOC
  h->d_nb = s.nbp;
  return ORCTM_NH_OK;
}

// Staging buffer for a Fortran (IE,JE,KE) array, allocated on first use.
hipError_t ensure_fort(orctm_nh_t* h) {
  if (h->d_fort) return hipSuccess;
  void* p = nullptr;
  hipError_t e = dalloc(h, &p, (size_t)h->IE * h->JE * h->KE * sizeof(double));
  if (e == hipSuccess) h->d_fort = static_cast<double*>(p);
  return e;
}

hipError_t ctx_build(Ctx& c, const std::vector<Sys>& systems, const std::vector<int>& forced,
                     const std::vector<int>& zw_table, const std::vector<int>& addressing, bool fp32, int check_every,
                     int max_outer, int device, bool prof, hipStream_t stream);
void ctx_destroy(Ctx& c);

// The single-system context of a handle, built on its stream on first use: a clone that is only ever
// solved inside a batch (the serving loop's requests) never builds one.
int ensure_ctx(orctm_nh_t* h) {
  if (h->ctx_ready) return ORCTM_NH_OK;
  DEVLOCK;
  hipError_t e = ctx_build(h->ctx, {h->sys}, {h->opts.tile_rows}, {h->opts.zw_table}, {h->opts.buffer_addressing},
                           h->fp32, h->opts.check_every, h->opts.max_outer, h->opts.device, h->opts.profile != 0,
                           h->stream);
  if (e != hipSuccess) {
    hipStreamSynchronize(h->stream);
    ctx_destroy(h->ctx);
    return map_err(e);
  }
  h->ctx_ready = true;
  return ORCTM_NH_OK;
}

void free_all(orctm_nh_t* h) {
  if (h->bot) bot_release(*h->bot);
  h->bot.reset();
  for (void* p : h->allocs) pool_free(h->opts.device, p);
  h->allocs.clear();
}

// Drains the handle's stream, returns its memory to the pool and deletes it.
void release(orctm_nh_t* h) {
  if (h->stream) hipStreamSynchronize(h->stream);
  ctx_destroy(h->ctx);
  free_all(h);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

// Tiles (64 x h columns) of one system in natural (row-major) order.
std::vector<Tile> system_tiles(int sys_index, int nx, int ny, int h) {
  std::vector<Tile> t;
  const int tx = (nx + TX - 1) / TX, ty = (ny + h - 1) / h;
  t.reserve((size_t)tx * ty);
  int local = 0;
  for (int by = 0; by < ty; ++by)
    for (int bx = 0; bx < tx; ++bx) t.push_back(Tile{sys_index, bx, by, local++});
  return t;
}

// The order of a streaming launch's tiles (nh_tiles.h): contiguous runs of equal work per XCD.
std::vector<Tile> balanced_order(const std::vector<Tile>& nat, const std::vector<Sys>& sys) {
  return onh::balanced_order(nat, [&](const Tile& t) { return sys[t.sys].nz; }, kDeepTiles, Tile{-1, 0, 0, 0});
}

// The caller has synchronised the context's stream (no queued work uses its blocks).
void ctx_destroy(Ctx& c) {
  for (auto& kv : c.graphs) hipGraphExecDestroy(kv.second);
  pool_free(c.device, c.d_sys);
  pool_free(c.device, c.d_tiles);
  pool_free(c.device, c.d_otiles);
  for (Group& g : c.groups) {
    pool_free(c.device, g.d_tiles);
    pool_free(c.device, g.d_pairs);
  }
  pool_free(c.device, c.d_outer);
  pool_free(c.device, c.d_inner);
  pool_free_host(c.h_outer);
  pool_free_host(c.h_inner);
  pool_free(c.device, c.d_seeds);
  pool_free_host(c.h_seeds);
  if (c.ev0) hipEventDestroy(c.ev0);
  if (c.ev1) hipEventDestroy(c.ev1);
  for (hipEvent_t e : c.kev) hipEventDestroy(e);
  if (c.stream && c.own_stream) hipStreamDestroy(c.stream);
  c = Ctx{};
}

// Host-to-device copy of a small host array on a stream (the host array may go out of scope once the
// stream has been synchronised).
hipError_t put_async(void* dst, const void* src, size_t bytes, hipStream_t st) {
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
}

// One inner-iteration phase over every tile group, then the phase's per-system reduction. With e0/e1
// given, events bracket the phase's kernels (not the reduction) for kernel timing.
hipError_t iter_bwd(Ctx& c, int raw, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (e0) HIPCHK(hipEventRecord(e0, c.stream));
  for (const Group& g : c.groups) HIPCHK(launch_bwd(c.stream, c.fp32, c.buf, g.h, c.d_sys, g.d_tiles, g.ntiles, raw));
  if (e1) HIPCHK(hipEventRecord(e1, c.stream));
  return launch_inner_reduce(c.stream, c.fp32, PH_BWD, c.d_sys, c.nsys);
}

hipError_t iter_fwd(Ctx& c, bool init, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (e0) HIPCHK(hipEventRecord(e0, c.stream));
  for (const Group& g : c.groups)
    HIPCHK(launch_fwd(c.stream, c.fp32, init, g.h, c.d_sys, g.d_tiles, g.ntiles, g.d_pairs, g.npairs));
  if (e1) HIPCHK(hipEventRecord(e1, c.stream));
  return launch_inner_reduce(c.stream, c.fp32, init ? PH_INIT : PH_FWD, c.d_sys, c.nsys);
}

// ---- tile-height tuning --------------------------------------------------------------------------
// The inner kernels run on 64 x h tiles, h in {4, 8, 12}. Small h gives more blocks (small systems
// fill the 256 CUs) and lighter blocks (two blocks per CU where registers are tight, FP64); large h
// re-reads fewer halo rows in the matvec. The best h is measured per size class (ceil(log2) of the
// column count nx*ny, the precision, and ceil(log2) of the number of systems of the class in the
// context) and kept for the life of the process.

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, int> g_tune;

// The choices persist in a JSON file, one entry per line, keyed by the device's architecture and CU
// count (a different device tunes afresh): ORCTM_NH_TUNE_FILE names it (empty: none), else
// $XDG_CACHE_HOME/orctm_nh/tile_tuning.json or ~/.cache/orctm_nh/tile_tuning.json. It is read once per
// device and process, before the first tuning, and rewritten whole (through a temporary file and a
// rename) after each new choice. Entries of another format version are ignored: the version changes
// whenever the kernels' relative speeds may have.
constexpr int kTuneFormat = 3;
std::map<int, bool> g_tune_loaded;  // device ordinal -> file read

std::string tune_path() {
  if (const char* e = getenv("ORCTM_NH_TUNE_FILE")) return std::string(e);
  if (const char* x = getenv("XDG_CACHE_HOME"))
    if (*x) return std::string(x) + "/orctm_nh/tile_tuning.json";
  if (const char* h = getenv("HOME"))
    if (*h) return std::string(h) + "/.cache/orctm_nh/tile_tuning.json";
  return std::string();
}

std::string device_identity(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return std::string();
  char buf[128];
  const char* colon = strchr(prop.gcnArchName, ':');
  const int n = colon ? (int)(colon - prop.gcnArchName) : (int)strlen(prop.gcnArchName);
  snprintf(buf, sizeof(buf), "%.*s/%d", n, prop.gcnArchName, prop.multiProcessorCount);
  return std::string(buf);
}

// Under g_tune_mu: entries of the file for this device into g_tune, once.
void tune_load(int device) {
  if (g_tune_loaded[device]) return;
  g_tune_loaded[device] = true;
  const std::string path = tune_path(), id = device_identity(device);
  if (path.empty() || id.empty()) return;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  char line[512], dev[128];
  int ver = 0, mode = 0, cls = 0, ncls = 0, rows = 0;
  while (fgets(line, sizeof(line), f)) {
    if (sscanf(line, " {\"format\": %d, \"device\": \"%127[^\"]\", \"mode\": %d, \"class\": %d, \"count_class\": %d, "
                     "\"rows\": %d}",
               &ver, dev, &mode, &cls, &ncls, &rows) != 6)
      continue;
    if (ver != kTuneFormat || id != dev) continue;
    const bool ok = mode >= 100 ? (rows == 0 || rows == 4 || rows == 8 || rows == 12) : (rows == 4 || rows == 8 || rows == 12);
    if (ok) g_tune.emplace(std::make_tuple(device, mode, cls, ncls), rows);
  }
  fclose(f);
}

// Under g_tune_mu: the file rewritten with every entry this process knows for this device, plus the
// file's entries for other devices and format versions, kept as they were.
void tune_save(int device) {
  const std::string path = tune_path(), id = device_identity(device);
  if (path.empty() || id.empty()) return;
  std::vector<std::string> keep;
  if (FILE* f = fopen(path.c_str(), "r")) {
    char line[512], dev[128];
    int ver = 0;
    while (fgets(line, sizeof(line), f))
      if (sscanf(line, " {\"format\": %d, \"device\": \"%127[^\"]\"", &ver, dev) == 2 &&
          (ver != kTuneFormat || id != dev)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == ',')) l.pop_back();
        keep.push_back(l);
      }
    fclose(f);
  }
  for (const auto& kv : g_tune) {
    if (std::get<0>(kv.first) != device) continue;
    char buf[256];
    snprintf(buf, sizeof(buf),
             "{\"format\": %d, \"device\": \"%s\", \"mode\": %d, \"class\": %d, \"count_class\": %d, \"rows\": %d}",
             kTuneFormat, id.c_str(), std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first), kv.second);
    keep.push_back(buf);
  }
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {  // create the directory chain (ignore failures: no cache then)
    for (size_t q = 1; q <= slash; ++q)
      if (path[q] == '/' || q == slash) mkdir(path.substr(0, q == slash ? slash : q).c_str(), 0755);
  }
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return;
  fprintf(f, "{\"entries\": [\n");
  for (size_t q = 0; q < keep.size(); ++q) fprintf(f, "%s%s\n", keep[q].c_str(), q + 1 < keep.size() ? "," : "");
  fprintf(f, "]}\n");
  const bool ok = fclose(f) == 0;
  if (ok) rename(tmp.c_str(), path.c_str());
  else remove(tmp.c_str());
}

int ceil_log2(long long v) {
  int c = 0;
  while ((1LL << c) < v) ++c;
  return c;
}

int size_class(const Sys& s) { return ceil_log2((long long)s.nx * s.ny); }

void order_into(std::vector<Tile>& nat, const Sys& s, int q, int h) {
  for (const Tile& t : system_tiles(q, s.nx, s.ny, h)) nat.push_back(t);
}

// Scratch copies of some systems of a context, for timing the inner kernels under a tile-height
// assignment. The kernels run a full iteration's memory traffic against scratch inner states (not
// the first iteration, alpha = beta = 0): they overwrite the inner work vectors w, p, q and the
// partial sums, all of which every solve re-initialises before use, and store x and r back
// unchanged; the FP64 solution, the right-hand side and the operator are not touched.
struct TuneSet {
  Ctx& c;
  std::vector<Sys> ts;
  Sys* d_ts = nullptr;
  void* d_in = nullptr;
  TuneSet(Ctx& ctx, const std::vector<Sys>& dev, const std::vector<int>& members) : c(ctx) {
    ts.resize(members.size());
    for (size_t q = 0; q < members.size(); ++q) ts[q] = dev[members[q]];
  }
  ~TuneSet() {
    hipStreamSynchronize(c.stream);
    pool_free(c.device, d_ts);
    pool_free(c.device, d_in);
  }
  hipError_t upload() {
    const size_t m = ts.size(), isz = inner_size(c.fp32);
    std::vector<char> hin(isz * m, 0);
    for (size_t q = 0; q < m; ++q) {
      if (c.fp32) {
        Inner<float>* in = reinterpret_cast<Inner<float>*>(hin.data()) + q;
        in->done = RUN;
        in->rho = 1.f;
      } else {
        Inner<double>* in = reinterpret_cast<Inner<double>*>(hin.data()) + q;
        in->done = RUN;
        in->rho = 1.0;
      }
    }
    void* p = nullptr;
    hipError_t e = pool_alloc(c.device, sizeof(Sys) * m, &p);
    d_ts = static_cast<Sys*>(p);
    if (e == hipSuccess) e = pool_alloc(c.device, hin.size(), &d_in);
    if (e == hipSuccess) e = put_async(d_in, hin.data(), hin.size(), c.stream);
    for (size_t q = 0; q < m; ++q) ts[q].inner = static_cast<char*>(d_in) + isz * q;
    if (e == hipSuccess) e = put_async(d_ts, ts.data(), sizeof(Sys) * m, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    return e;
  }
  // one warm round, then 3 timed rounds of (matvec, column update), heights h[q] per system; the
  // fastest round counts, so that kernels of other streams on the device (serving lanes) sway the
  // choice less
  hipError_t time(const std::vector<int>& h, float* ms) {
    std::vector<Tile*> d_tiles, d_pairs;
    std::vector<int> nts, nps, hs;
    hipError_t e = hipSuccess;
    for (int hc : kTileHeights) {
      std::vector<Tile> nat;
      for (size_t q = 0; q < ts.size(); ++q)
        if (h[q] == hc) order_into(nat, ts[q], (int)q, hc);
      if (nat.empty()) continue;
      std::vector<Tile> ord = balanced_order(nat, ts), prs = balanced_order(pair_tiles(nat), ts);
      void *d = nullptr, *dp = nullptr;
      e = pool_alloc(c.device, sizeof(Tile) * ord.size(), &d);
      if (e != hipSuccess) break;
      d_tiles.push_back(static_cast<Tile*>(d));
      e = pool_alloc(c.device, sizeof(Tile) * std::max<size_t>(prs.size(), 1), &dp);
      if (e != hipSuccess) break;
      d_pairs.push_back(static_cast<Tile*>(dp));
      nts.push_back((int)ord.size());
      nps.push_back((int)prs.size());
      hs.push_back(hc);
      e = put_async(d, ord.data(), sizeof(Tile) * ord.size(), c.stream);
      if (e == hipSuccess && !prs.empty()) e = put_async(dp, prs.data(), sizeof(Tile) * prs.size(), c.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
      if (e != hipSuccess) break;
    }
    float best = -1.f;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
      if (rep > 0) e = hipEventRecord(c.ev0, c.stream);
      for (size_t g = 0; g < hs.size() && e == hipSuccess; ++g)
        e = launch_bwd(c.stream, c.fp32, c.buf, hs[g], d_ts, d_tiles[g], nts[g], 0);
      for (size_t g = 0; g < hs.size() && e == hipSuccess; ++g)
        e = launch_fwd(c.stream, c.fp32, false, hs[g], d_ts, d_tiles[g], nts[g], d_pairs[g], nps[g]);
      if (rep == 0 || e != hipSuccess) continue;
      float t = 0.f;
      e = hipEventRecord(c.ev1, c.stream);
      if (e == hipSuccess) e = hipEventSynchronize(c.ev1);
      if (e == hipSuccess) e = hipEventElapsedTime(&t, c.ev0, c.ev1);
      if (e == hipSuccess && (best < 0.f || t < best)) best = t;
    }
    *ms = best;
    hipStreamSynchronize(c.stream);
    for (Tile* d : d_tiles) pool_free(c.device, d);
    for (Tile* d : d_pairs) pool_free(c.device, d);
    return e;
  }
};

// Chooses each system's tile height and builds the iteration tile groups, one per height in use,
// each in XCD order. forced[q] > 0 fixes a system's height. Systems (or contexts) large enough to
// fill the device with 64 x 12 tiles take 12 rows untimed. The others are tuned in two stages:
// (1) per size class, timing the class's systems alone at each height; (2) for a context of several
// systems, the whole launch is timed under the per-class choice and under each uniform height (all
// systems run together, so a class that prefers small tiles alone may not inside a large batch),
// and the fastest assignment is kept. Both choices are cached for the life of the process and in
// the tuning file (tune_path), keyed by device, precision, size class (ceil(log2) of the column
// count; of the total for stage 2) and ceil(log2) of the number of systems.
// Tiles of 64 x 12 needed before a set of systems fills the device by itself (8 blocks per CU):
// beyond it the tallest tile (fewest halo re-reads) is the fastest and no timing is run.
constexpr long long kFillTiles = 8 * 256;

long long tiles12(const std::vector<Sys>& dev, const std::vector<int>& members) {
  long long n = 0;
  for (int q : members) n += (long long)((dev[q].nx + TX - 1) / TX) * ((dev[q].ny + 11) / 12);
  return n;
}

hipError_t ctx_heights(Ctx& c, const std::vector<Sys>& dev, const std::vector<int>& forced) {
  c.heights.assign(c.nsys, 12);
  std::map<int, std::vector<int>> classes;
  std::vector<int> tuned;
  long long cols = 0;
  for (int q = 0; q < c.nsys; ++q) {
    if (forced[q] > 0) {
      c.heights[q] = forced[q];
    } else {
      classes[size_class(dev[q])].push_back(q);
      tuned.push_back(q);
      cols += (long long)dev[q].nx * dev[q].ny;
    }
  }
  for (auto& kv : classes) {
    const auto key = std::make_tuple(c.device, c.fp32 ? 1 : 0, kv.first, ceil_log2((long long)kv.second.size()));
    int h = 0;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tune_load(c.device);
      auto it = g_tune.find(key);
      if (it != g_tune.end()) h = it->second;
    }
    if (!h && tiles12(dev, kv.second) >= kFillTiles) h = 12;
    if (!h) {
      TuneSet ts(c, dev, kv.second);
      HIPCHK(ts.upload());
      float best = -1.f;
      for (int hc : kTileHeights) {
        float ms = 0.f;
        HIPCHK(ts.time(std::vector<int>(kv.second.size(), hc), &ms));
        if (best < 0.f || ms < best) {
          best = ms;
          h = hc;
        }
      }
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[key] = h;
      tune_save(c.device);
    }
    for (int q : kv.second) c.heights[q] = h;
  }
  if (tuned.size() > 1 && classes.size() > 0) {
    // stage 2: the whole launch; code 0 = keep the per-class choice, else a uniform height
    const auto key = std::make_tuple(c.device, 100 + (c.fp32 ? 1 : 0), ceil_log2(cols), ceil_log2((long long)tuned.size()));
    int code = -1;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tune_load(c.device);
      auto it = g_tune.find(key);
      if (it != g_tune.end()) code = it->second;
    }
    if (code < 0 && tiles12(dev, tuned) >= kFillTiles) code = 12;
    if (code < 0) {
      TuneSet ts(c, dev, tuned);
      HIPCHK(ts.upload());
      std::vector<int> mix(tuned.size());
      for (size_t q = 0; q < tuned.size(); ++q) mix[q] = c.heights[tuned[q]];
      float best = 0.f;
      HIPCHK(ts.time(mix, &best));
      code = 0;
      for (int hc : kTileHeights) {
        float ms = 0.f;
        HIPCHK(ts.time(std::vector<int>(tuned.size(), hc), &ms));
        if (ms < best) {
          best = ms;
          code = hc;
        }
      }
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[key] = code;
      tune_save(c.device);
    }
    if (code > 0)
      for (int q : tuned) c.heights[q] = code;
  }
  for (int hc : kTileHeights) {
    std::vector<Tile> nat;
    for (int q = 0; q < c.nsys; ++q)
      if (c.heights[q] == hc) order_into(nat, dev[q], q, hc);
    if (nat.empty()) continue;
    std::vector<Tile> ord = balanced_order(nat, dev), prs = balanced_order(pair_tiles(nat), dev);
    c.groups.push_back(Group{});
    Group& g = c.groups.back();
    g.h = hc;
    g.ntiles = (int)ord.size();
    g.npairs = (int)prs.size();
    void* p = nullptr;
    HIPCHK(pool_alloc(c.device, sizeof(Tile) * g.ntiles, &p));
    g.d_tiles = static_cast<Tile*>(p);
    HIPCHK(pool_alloc(c.device, sizeof(Tile) * std::max(g.npairs, 1), &p));
    g.d_pairs = static_cast<Tile*>(p);
    HIPCHK(put_async(g.d_tiles, ord.data(), sizeof(Tile) * g.ntiles, c.stream));
    if (g.npairs) HIPCHK(put_async(g.d_pairs, prs.data(), sizeof(Tile) * g.npairs, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
  }
  return hipSuccess;
}

// Builds a context over the given systems (host Sys copies; outer/inner pointers are replaced by
// the context's own contiguous arrays). forced[q]: tile height of system q, 0 = tuned.
// Solve contexts of at least this many cells read ZW/ZP from the table by default (options.zw_table 0).
// Measured on gfx950 with the table forced on and off: it raises the rate of a 1024x768x60 system by
// 12% and lowers that of a 512x4x100 transect by 14%, whose launches fill a few dozen CUs.
constexpr long long kTabMinCells = 1LL << 20;

// stream: the owning handle's stream for a single-system context, or nullptr for a context of its own
// (a batch), which then creates one. Everything the build does is ordered on that stream.
hipError_t ctx_build(Ctx& c, const std::vector<Sys>& systems, const std::vector<int>& forced,
                     const std::vector<int>& zw_table, const std::vector<int>& addressing, bool fp32, int check_every,
                     int max_outer, int device, bool prof, hipStream_t stream) {
  c.nsys = (int)systems.size();
  c.buf = true;
  for (int q = 0; q < c.nsys; ++q) c.buf = c.buf && addressing[q] != -1 && bwd_buffers_fit(systems[q], fp32);
  c.prof = prof;
  c.iter_seen.assign(c.nsys, 0);
  c.lrate.assign(c.nsys, 0.0);
  for (const Sys& s : systems) c.cells += (long long)s.nx * s.ny * s.nz;
  c.fp32 = fp32;
  c.check_every = check_every > 0 ? check_every : 16;
  c.max_outer = max_outer;
  c.device = device;
  for (const Sys& s : systems) c.maxit = std::max(c.maxit, s.maxit);
  std::vector<Tile> nat, onat;
  for (int q = 0; q < c.nsys; ++q) order_into(nat, systems[q], q, TY);
  for (int q = 0; q < c.nsys; ++q) order_into(onat, systems[q], q, kOuterRows);
  std::vector<Tile> ord = xcd_order(nat), oord = balanced_order(onat, systems);
  c.ntiles = (int)ord.size();
  c.n_otiles = (int)oord.size();
  if (stream) {
    c.stream = stream;
    c.own_stream = false;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  }
  if (c.prof) {
    c.kev.assign(4 * (size_t)c.check_every, nullptr);
    for (hipEvent_t& e : c.kev) HIPCHK(hipEventCreate(&e));
  }
  HIPCHK(hipEventCreate(&c.ev0));
  HIPCHK(hipEventCreate(&c.ev1));
  void* p = nullptr;
  HIPCHK(pool_alloc(device, sizeof(Outer) * c.nsys, &p));
  c.d_outer = static_cast<Outer*>(p);
  HIPCHK(hipMemsetAsync(c.d_outer, 0, sizeof(Outer) * c.nsys, c.stream));
  HIPCHK(pool_alloc(device, inner_size(fp32) * c.nsys, &c.d_inner));
  HIPCHK(hipMemsetAsync(c.d_inner, 0, inner_size(fp32) * c.nsys, c.stream));
  HIPCHK(pool_alloc_host(sizeof(Outer) * c.nsys, &p));
  c.h_outer = static_cast<Outer*>(p);
  HIPCHK(pool_alloc_host(inner_size(fp32) * c.nsys, &c.h_inner));
  HIPCHK(pool_alloc(device, sizeof(unsigned long long) * c.nsys, &p));
  c.d_seeds = static_cast<unsigned long long*>(p);
  HIPCHK(pool_alloc_host(sizeof(unsigned long long) * c.nsys, &p));
  c.h_seeds = static_cast<unsigned long long*>(p);
  std::vector<Sys> dev = systems;
  c.tab.assign(c.nsys, 0);
  for (int q = 0; q < c.nsys; ++q) {
    dev[q].outer = c.d_outer + q;
    dev[q].inner = static_cast<char*>(c.d_inner) + inner_size(fp32) * q;
    const int mode = zw_table[q];
    dev[q].tab = dev[q].tab && (mode == 1 || (mode == 0 && c.cells >= kTabMinCells)) ? 1 : 0;
    c.tab[q] = dev[q].tab;
  }
  HIPCHK(pool_alloc(device, sizeof(Sys) * c.nsys, &p));
  c.d_sys = static_cast<Sys*>(p);
  HIPCHK(put_async(c.d_sys, dev.data(), sizeof(Sys) * c.nsys, c.stream));
  HIPCHK(pool_alloc(device, sizeof(Tile) * std::max(c.ntiles, 1), &p));
  c.d_tiles = static_cast<Tile*>(p);
  if (c.ntiles) HIPCHK(put_async(c.d_tiles, ord.data(), sizeof(Tile) * c.ntiles, c.stream));
  HIPCHK(pool_alloc(device, sizeof(Tile) * std::max(c.n_otiles, 1), &p));
  c.d_otiles = static_cast<Tile*>(p);
  if (c.n_otiles) HIPCHK(put_async(c.d_otiles, oord.data(), sizeof(Tile) * c.n_otiles, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return ctx_heights(c, dev, forced);
}

// n inner iterations (matvec + column update each). Runs of a power-of-two length are captured
// once into a hipGraph and replayed; n is issued as its binary decomposition, largest run first.
// With kernel timing on, the same launches are issued directly with HIP events around each phase.
hipError_t ctx_chunk(Ctx& c, int n) {
  if (c.prof) {
    for (int it = 0; it < n; ++it) {
      HIPCHK(iter_bwd(c, 0, c.kev[4 * it], c.kev[4 * it + 1]));
      HIPCHK(iter_fwd(c, false, c.kev[4 * it + 2], c.kev[4 * it + 3]));
    }
    return hipSuccess;
  }
  if (!c.use_graphs) {
    for (int it = 0; it < n; ++it) {
      HIPCHK(iter_bwd(c, 0));
      HIPCHK(iter_fwd(c, false));
    }
    return hipSuccess;
  }
  for (int len = 1 << 30; len > 0; len >>= 1) {
    if (!(n & len)) continue;
    hipGraphExec_t& ge = c.graphs[len];
    if (!ge) {
      DEVLOCK;
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      for (int it = 0; it < len && e == hipSuccess; ++it) {
        e = iter_bwd(c, 0);
        if (e == hipSuccess) e = iter_fwd(c, false);
      }
      hipError_t e2 = hipStreamEndCapture(c.stream, &g);
      if (e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        c.graphs.erase(len);
        return e;
      }
      if (e2 != hipSuccess) c.graphs.erase(len);
      HIPCHK(e2);
      e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) c.graphs.erase(len);
      HIPCHK(e);
    }
    HIPCHK(hipGraphLaunch(ge, c.stream));
  }
  return hipSuccess;
}

// Inner iterations to issue before the next poll: enough for every running system to reach its
// tolerance if its squared residual keeps falling at its mean rate (log(rr/rr0)/iter, from this
// inner solve once it has run, else from the system's previous inner solve), at least 1 and at most
// check_every. Launches past a system's convergence exit at entry, but each still costs a few
// microseconds, and a run of them drains the queue faster than the runtime submits the graph's
// nodes (a stall of about a millisecond), so the chunk is sized to the work left. The first chunk
// of an inner solve uses the tolerance the reduction kernel sets, max(rtol_inner, rtol/2 |b|/|r|)^2,
// computed here from the outer state read at the start of the refinement step.
int chunk_length(Ctx& c, bool first) {
  int need = 1;
  for (int q = 0; q < c.nsys; ++q) {
    int n = c.check_every;
    if (first) {
      const Outer& o = c.h_outer[q];
      if (o.odone != RUN) continue;
      double tol = 0.5 * o.rtol * o.nb / o.nr;
      if (tol < o.rtol_inner) tol = o.rtol_inner;
      if (c.lrate[q] < 0.0 && tol > 0.0 && tol < 1.0) n = (int)std::ceil(2.0 * std::log(tol) / c.lrate[q]);
    } else {
      if (inner_done(c, q) != RUN) continue;
      double rr, rr0, tol2;
      if (c.fp32) {
        const Inner<float>& in = reinterpret_cast<const Inner<float>*>(c.h_inner)[q];
        rr = in.rr, rr0 = in.rr0, tol2 = in.tol2;
      } else {
        const Inner<double>& in = reinterpret_cast<const Inner<double>*>(c.h_inner)[q];
        rr = in.rr, rr0 = in.rr0, tol2 = in.tol2;
      }
      const int it = inner_iter(c, q);
      const double red = rr / rr0;
      if (it > 0 && red > 0.0 && red < 1.0 && tol2 > 0.0 && std::isfinite(red)) {
        c.lrate[q] = std::log(red) / it;
        const double left = std::log(tol2 / red) / c.lrate[q];
        n = left <= 1.0 ? 1 : (int)std::ceil(left);
      }
    }
    need = std::max(need, std::min(std::max(n, 1), c.check_every));
  }
  return need;
}

// Runs the whole refinement for every system of the context; results stay on the device.
hipError_t ctx_solve(Ctx& c, float* ms) {
  HIPCHK(hipEventRecord(c.ev0, c.stream));
  HIPCHK(launch_begin(c.stream, c.fp32, c.d_sys, c.nsys));
  HIPCHK(launch_residual(c.stream, c.fp32, kOuterRows, c.d_sys, c.d_otiles, c.n_otiles, c.nsys));
  for (int m = 0; m <= c.max_outer + 1; ++m) {
    HIPCHK(hipMemcpyAsync(c.h_outer, c.d_outer, sizeof(Outer) * c.nsys, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    bool any = false;
    for (int q = 0; q < c.nsys; ++q) any |= (c.h_outer[q].odone == RUN);
    if (!any) break;
    HIPCHK(iter_fwd(c, true));
    std::fill(c.iter_seen.begin(), c.iter_seen.end(), 0);
    // every inner solve ends within maxit iterations (device-side NOTCONV); the host bound on the
    // iterations issued only guarantees termination of this loop whatever the device flags say
    const long long max_issue = (long long)c.maxit + 2LL * c.check_every;
    long long issued = 0;
    for (int chunk = 0; issued < max_issue; ++chunk) {
      const int n = chunk_length(c, chunk == 0);
      issued += n;
      HIPCHK(ctx_chunk(c, n));
      HIPCHK(hipMemcpyAsync(c.h_inner, c.d_inner, inner_size(c.fp32) * c.nsys, hipMemcpyDeviceToHost, c.stream));
      HIPCHK(hipStreamSynchronize(c.stream));
      bool run = false;
      int worked = 0;
      for (int q = 0; q < c.nsys; ++q) {
        run |= (inner_done(c, q) == RUN);
        const int it = inner_iter(c, q);
        worked = std::max(worked, it - c.iter_seen[q]);
        c.iter_seen[q] = it;
        double rr, rr0;
        if (c.fp32) {
          const Inner<float>& in = reinterpret_cast<const Inner<float>*>(c.h_inner)[q];
          rr = in.rr, rr0 = in.rr0;
        } else {
          const Inner<double>& in = reinterpret_cast<const Inner<double>*>(c.h_inner)[q];
          rr = in.rr, rr0 = in.rr0;
        }
        const double red = rr / rr0;
        if (it > 0 && red > 0.0 && red < 1.0 && std::isfinite(red)) c.lrate[q] = std::log(red) / it;
      }
      if (c.prof) {
        for (int it = 0; it < std::min(worked, n); ++it) {
          float a = 0.f, b = 0.f;
          HIPCHK(hipEventElapsedTime(&a, c.kev[4 * it], c.kev[4 * it + 1]));
          HIPCHK(hipEventElapsedTime(&b, c.kev[4 * it + 2], c.kev[4 * it + 3]));
          c.mv_ms += a;
          c.up_ms += b;
          c.mv_n += 1;
          c.up_n += 1;
        }
      }
      if (!run) break;
    }
    HIPCHK(launch_refine(c.stream, c.fp32, TY, c.d_sys, c.d_tiles, c.ntiles));
    HIPCHK(launch_residual(c.stream, c.fp32, kOuterRows, c.d_sys, c.d_otiles, c.n_otiles, c.nsys));
  }
  HIPCHK(hipEventRecord(c.ev1, c.stream));
  HIPCHK(hipMemcpyAsync(c.h_outer, c.d_outer, sizeof(Outer) * c.nsys, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  if (ms) HIPCHK(hipEventElapsedTime(ms, c.ev0, c.ev1));
  return hipSuccess;
}

// b <- b (1 + amplitude xi) for every system of the context, system q with its own seed; ||b|| is
// refreshed. Synchronous.
hipError_t ctx_perturb(Ctx& c, double amplitude, const uint64_t* seeds) {
  for (int q = 0; q < c.nsys; ++q) c.h_seeds[q] = seeds[q];
  HIPCHK(hipMemcpyAsync(c.d_seeds, c.h_seeds, sizeof(unsigned long long) * c.nsys, hipMemcpyHostToDevice, c.stream));
  HIPCHK(launch_perturb(c.stream, c.d_sys, c.d_tiles, c.ntiles, c.nsys, amplitude, c.d_seeds));
  return hipStreamSynchronize(c.stream);
}

int outer_status(const Outer& o) {
  switch (o.odone) {
    case CONVERGED: return ORCTM_NH_OK;
    case NOTCONV: return ORCTM_NH_NOT_CONVERGED;
    case NONFINITE: return ORCTM_NH_ERR_NONFINITE;
    case RUN: return ORCTM_NH_NOT_CONVERGED;
    default: return ORCTM_NH_BREAKDOWN;
  }
}

void fill_stats(const Outer& o, int tile_rows, int zw_table, bool buf, orctm_nh_stats_t* st, double solve_ms,
                double total_ms) {
  if (!st) return;
  memset(st, 0, sizeof(*st));
  st->status = outer_status(o);
  st->tile_rows = tile_rows;
  st->zw_table = zw_table;
  st->buffer_addressing = buf ? 1 : 0;
  st->outer_iters = o.outer;
  st->inner_iters = o.inner_total;
  st->rel_res = o.rel;
  st->bnorm = o.nb;
  st->solve_ms = solve_ms;
  st->total_ms = total_ms;
}

bool grid_valid(const orctm_nh_grid_t* g) {
  if (!g) return false;
  if (g->ie < 3 || g->je < 3 || g->ke < 1) return false;
  if (g->i_start < 0 || g->i_start > 1 || g->j_start < 0 || g->j_start > 1) return false;
  if (!g->dlxp || !g->dlyp || !g->dlxu || !g->dlyv || !g->ddpo || !g->weto || !g->dzw) return false;
  const long long cells = (long long)(g->ie - 2) * (g->je - 2) * g->ke;
  return cells > 0 && cells < (1LL << 40);
}

hipError_t upload(void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); }

// The 2-D bottom-pressure solve's device state (options.bottom_solve): the metrics it needs from the
// grid, the zeroed module arrays (beleg_zero.f90:481-504), the work vectors and the 2-D operator.
hipError_t bot_create(orctm_nh_t* h, const orctm_nh_grid_t* g) {
  auto b = std::make_unique<BotState>();
  b->IE = h->IE;
  b->JE = h->JE;
  b->KE = h->KE;
  b->IS = h->i_start;
  b->JS = h->j_start;
  b->cyc_x = h->cyc_x;
  b->cyc_y = h->cyc_y;
  b->nx = h->nx;
  b->ny = h->ny;
  const size_t IE = h->IE, JE = h->JE;
  const size_t p2 = IE * JE, u2 = (IE - h->i_start + 1) * JE, v2 = IE * (JE - h->j_start + 1);
  const size_t n = (size_t)h->nx * h->ny, nblk = (n + 255) / 256;
  hipError_t e = hipSuccess;
  auto buf = [&](double*& dst, size_t cnt) {
    void* p = nullptr;
    if (e == hipSuccess) e = dalloc(h, &p, cnt * sizeof(double));
    dst = static_cast<double*>(p);
  };
  auto put = [&](const double*& dst, const double* src, size_t cnt) {
    double* p = nullptr;
    buf(p, cnt);
    if (e == hipSuccess) e = upload(p, src, cnt * sizeof(double));
    dst = p;
  };
  put(b->dlxu, g->dlxu, u2);
  put(b->dlyv, g->dlyv, v2);
  put(b->dlxp, g->dlxp, p2);
  put(b->dlyp, g->dlyp, p2);
  put(b->dzw, g->dzw, h->KE);
  buf(b->coef, 5 * n);
  buf(b->S, n);
  buf(b->botgu, u2);
  buf(b->botgv, v2);
  buf(b->botgw, p2);
  buf(b->dbs, p2);
  buf(b->pnh_bot, p2);
  buf(b->b, n);
  buf(b->x, n);
  buf(b->r, n);
  buf(b->z, n);
  buf(b->d0, n);
  buf(b->d1, n);
  buf(b->q, n);
  buf(b->part, 3 * nblk);
  buf(b->scal, 8);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->scal_host, 8 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = bot_setup(h->stream, *b);
  if (e != hipSuccess) {
    bot_release(*b);
    return e;
  }
  h->bot = std::move(b);
  return hipSuccess;
}

// Device part of create: allocations, assembly, factorisation, single-system context.
int create_device(orctm_nh_t* h, const orctm_nh_grid_t* g) {
  const size_t NP = (size_t)h->plane * h->nz;
  const size_t es = h->fp32 ? sizeof(float) : sizeof(double);
  Sys& s = h->sys;
  memset(&s, 0, sizeof(s));
  s.nx = h->nx;
  s.ny = h->ny;
  s.nz = h->nz;
  s.pitch = h->pitch;
  s.plane = h->plane;
  s.cyc_x = g->cyclic_x ? 1 : 0;
  s.cyc_y = g->cyclic_y ? 1 : 0;
  s.tiles_x = (h->nx + TX - 1) / TX;
  s.tiles_y = (h->ny + TY - 1) / TY;
  s.ntiles = s.tiles_x * s.tiles_y;
  s.nslots = s.tiles_x * ((h->ny + RG - 1) / RG);
  void* p;
  hipError_t e;
#define ALLOC(field, bytes, type)              \
  do {                                         \
    e = opalloc(h, &p, (bytes));               \
    if (e != hipSuccess) return map_err(e);    \
    s.field = reinterpret_cast<type>(p);       \
  } while (0)
  const size_t NP1 = (size_t)h->plane * (h->nz + 1);
  ALLOC(zw, NP1 * sizeof(double), const double*);
  ALLOC(zp, NP * sizeof(double), const double*);
  ALLOC(flg, NP * sizeof(uint16_t), const uint16_t*);
  ALLOC(col, 7 * (size_t)h->plane * sizeof(double), const double*);
  ALLOC(tw, (size_t)(h->nz + 1) * sizeof(double), const double*);
  ALLOC(tp, (size_t)h->nz * sizeof(double), const double*);
  ALLOC(xk, (size_t)h->plane * sizeof(int), const int*);
  ALLOC(xf, 3 * (size_t)h->plane * sizeof(uint32_t), const uint32_t*);
  ALLOC(xw, 3 * (size_t)h->plane * sizeof(double), const double*);
  ALLOC(xp, 3 * (size_t)h->plane * sizeof(double), const double*);
  if (h->fp32) {
    ALLOC(zw_in, NP1 * sizeof(float), const void*);
    ALLOC(zp_in, NP * sizeof(float), const void*);
    ALLOC(col_in, 7 * (size_t)h->plane * sizeof(float), const void*);
    ALLOC(tw_in, (size_t)(h->nz + 1) * sizeof(float), const void*);
    ALLOC(tp_in, (size_t)h->nz * sizeof(float), const void*);
    ALLOC(xw_in, 3 * (size_t)h->plane * sizeof(float), const void*);
    ALLOC(xp_in, 3 * (size_t)h->plane * sizeof(float), const void*);
  } else {
    s.zw_in = s.zw;
    s.zp_in = s.zp;
    s.col_in = s.col;
    s.tw_in = s.tw;
    s.tp_in = s.tp;
    s.xw_in = s.xw;
    s.xp_in = s.xp;
  }
  s.tab = 0;
  // line-preconditioner factors: fp16 bits + a per-column scale for FP32, FP64 values otherwise
  const size_t fs = h->fp32 ? sizeof(uint16_t) : sizeof(double);
  ALLOC(invD, NP * fs, const void*);
  ALLOC(L, NP * fs, const void*);
  if (h->fp32) ALLOC(dscale, (size_t)h->plane * sizeof(float), const float*);
  ALLOC(smaxp, sizeof(unsigned long long), const double*);
#undef ALLOC
  h->d_smax = reinterpret_cast<unsigned long long*>(const_cast<double*>(s.smaxp));
  int arc = alloc_state(h);
  if (arc != ORCTM_NH_OK) return arc;
  s.rtol = h->opts.rtol;
  s.rtol_inner = h->opts.rtol_inner;
  s.maxit = h->opts.maxit;
  s.max_outer = h->opts.max_outer;

  // temporary device copies of the Fortran inputs
  const size_t n2 = (size_t)g->ie * g->je, n3 = n2 * g->ke;
  const size_t nu = (size_t)(g->ie - g->i_start + 1) * g->je, nv = (size_t)g->ie * (g->je - g->j_start + 1);
  // temporaries from the pool, all work ordered on the handle's stream
  const int dev = h->opts.device;
  hipStream_t st = h->stream;
  double *d_dlxp = nullptr, *d_dlyp = nullptr, *d_dlxu = nullptr, *d_dlyv = nullptr, *d_ddpo = nullptr,
         *d_weto = nullptr, *d_dzw = nullptr;
  Sys* d_sys1 = nullptr;
  Tile* d_tiles1 = nullptr;
  int* d_bad = nullptr;
  int* d_zbad = nullptr;
  int bad = 0, zbad = 1, rc = ORCTM_NH_OK;
  std::vector<Tile> tl = system_tiles(0, h->nx, h->ny, TY);
  auto tmp = [&](auto*& ptr, size_t bytes) {
    void* q = nullptr;
    if (e == hipSuccess) e = pool_alloc(dev, bytes, &q);
    ptr = static_cast<std::remove_reference_t<decltype(ptr)>>(q);
  };
  auto put = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
  };
  e = hipSuccess;
  tmp(d_dlxp, n2 * 8);
  tmp(d_dlyp, n2 * 8);
  tmp(d_dlxu, nu * 8);
  tmp(d_dlyv, nv * 8);
  tmp(d_ddpo, n3 * 8);
  tmp(d_weto, n3 * 8);
  tmp(d_dzw, (size_t)g->ke * 8);
  tmp(d_sys1, sizeof(Sys));
  tmp(d_tiles1, sizeof(Tile) * tl.size());
  tmp(d_bad, sizeof(int));
  tmp(d_zbad, sizeof(int));
  put(d_dlxp, g->dlxp, n2 * 8);
  put(d_dlyp, g->dlyp, n2 * 8);
  put(d_dlxu, g->dlxu, nu * 8);
  put(d_dlyv, g->dlyv, nv * 8);
  put(d_ddpo, g->ddpo, n3 * 8);
  put(d_weto, g->weto, n3 * 8);
  put(d_dzw, g->dzw, (size_t)g->ke * 8);
  put(d_sys1, &s, sizeof(Sys));
  put(d_tiles1, tl.data(), sizeof(Tile) * tl.size());
  if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, sizeof(int), st);
  if (e == hipSuccess) e = hipMemsetAsync(d_zbad, 0, sizeof(int), st);
  if (e == hipSuccess) {
    FortranIn f{};
    f.IE = g->ie;
    f.JE = g->je;
    f.KE = g->ke;
    f.i_start = g->i_start;
    f.j_start = g->j_start;
    f.cyc_x = g->cyclic_x ? 1 : 0;
    f.cyc_y = g->cyclic_y ? 1 : 0;
    f.open_w = g->open_w;
    f.open_e = g->open_e;
    f.open_n = g->open_n;
    f.open_s = g->open_s;
    f.dlxp = d_dlxp;
    f.dlyp = d_dlyp;
    f.dlxu = d_dlxu;
    f.dlyv = d_dlyv;
    f.ddpo = d_ddpo;
    f.weto = d_weto;
    f.dzw = d_dzw;
    e = launch_assemble(st, h->fp32, d_sys1, d_tiles1, (int)tl.size(), f, d_bad, h->d_smax);
    if (e == hipSuccess) e = launch_line_factor(st, h->fp32, d_sys1, d_tiles1, (int)tl.size());
    if (e == hipSuccess) e = launch_zwindow(st, h->fp32, d_sys1, d_tiles1, (int)tl.size(), f, d_zbad);
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&zbad, d_zbad, sizeof(int), hipMemcpyDeviceToHost, st);
  }
  // every temporary is idle once the stream has drained (also after a failure part-way)
  const hipError_t esync = hipStreamSynchronize(st);
  if (e == hipSuccess) e = esync;
  if (e != hipSuccess) rc = map_err(e);
  else if (bad) rc = ORCTM_NH_ERR_INVALID;
  for (void* q : {(void*)d_dlxp, (void*)d_dlyp, (void*)d_dlxu, (void*)d_dlyv, (void*)d_ddpo, (void*)d_weto,
                  (void*)d_dzw, (void*)d_sys1, (void*)d_tiles1, (void*)d_bad, (void*)d_zbad})
    pool_free(dev, q);
  if (rc != ORCTM_NH_OK) return rc;
  s.tab = zbad ? 0 : 1;
  return ensure_ctx(h);
}

// Padded device vector <-> natural-order host vector (elem bytes each).
hipError_t to_device_vec(const orctm_nh_t* h, void* d, const void* host, size_t es) {
  return hipMemcpy2D(d, (size_t)h->pitch * es, host, (size_t)h->nx * es, (size_t)h->nx * es,
                     (size_t)h->ny * h->nz, hipMemcpyHostToDevice);
}
hipError_t to_host_vec(const orctm_nh_t* h, void* host, const void* d, size_t es) {
  return hipMemcpy2D(host, (size_t)h->nx * es, d, (size_t)h->pitch * es, (size_t)h->nx * es, (size_t)h->ny * h->nz,
                     hipMemcpyDeviceToHost);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Staging of orctm_nh_apply_noalloc, for callers that cannot allocate device memory: x and y of at
// most kScratchCells padded cells each, in one file-scope device array. Every GPU the library runs
// on holds its own copy (128 MiB) from the loading of the code object until the process ends. The
// array's name is a host-side stand-in: host code uses only the address hipGetSymbolAddress returns
// for the current device, looked up once per device and kept in g_scratch_ptr.
constexpr size_t kScratchCells = size_t(1) << 23;
static __device__ __align__(256) double scratch[2 * kScratchCells];
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

hipError_t scratch_address(int device, double** out) {
  *out = nullptr;
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    DEVLOCK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

int prepare_guess(orctm_nh_t* h) {
  // warm start: keep p64 (previous solution or set_guess); otherwise start from zero
  if (!h->opts.warm_start && !h->guess_set) {
    hipError_t e = hipMemsetAsync(h->sys.p64, 0, (size_t)h->plane * h->nz * sizeof(double), h->stream);
    if (e != hipSuccess) return map_err(e);
  }
  h->guess_set = false;
  return ORCTM_NH_OK;
}

}  // namespace

extern "C" {

int orctm_nh_abi_version(void) { return ORCTM_NH_ABI_VERSION; }

const char* orctm_nh_last_error(void) { return g_last_error; }

int orctm_nh_device_count(int32_t* n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (n) *n = (e == hipSuccess) ? c : 0;
  return e == hipSuccess ? ORCTM_NH_OK : ORCTM_NH_ERR_NODEVICE;
}

void orctm_nh_default_options(orctm_nh_options_t* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->precision = ORCTM_NH_MIXED;
  o->maxit = 10000;
  o->max_outer = 40;
  o->warm_start = 1;
  o->rtol = 1e-10;
  o->rtol_inner = -1.0;  // resolved at create: 1e-5 (MIXED) or 0 (FP64)
  o->check_every = 16;
  o->device = 0;
}

const char* orctm_nh_status_string(int s) {
  switch (s) {
    case ORCTM_NH_OK: return "converged";
    case ORCTM_NH_NOT_CONVERGED: return "not converged (maxit/max_outer reached)";
    case ORCTM_NH_BREAKDOWN: return "Krylov breakdown";
    case ORCTM_NH_ERR_NONFINITE: return "non-finite right-hand side";
    case ORCTM_NH_ERR_INVALID: return "invalid argument or geometry";
    case ORCTM_NH_ERR_NOMEM: return "out of device memory";
    case ORCTM_NH_ERR_HIP: return "HIP runtime error";
    case ORCTM_NH_ERR_NODEVICE: return "no usable gfx950 device";
    case ORCTM_NH_PENDING: return "pending";
    default: return "unknown status";
  }
}

int orctm_nh_create(const orctm_nh_grid_t* grid, const orctm_nh_options_t* opts, orctm_nh_t** handle) {
  if (!handle) return ORCTM_NH_ERR_INVALID;
  DEVLOCK;
  *handle = nullptr;
  if (!grid_valid(grid)) return ORCTM_NH_ERR_INVALID;
  orctm_nh_options_t o;
  orctm_nh_default_options(&o);
  if (opts) o = *opts;
  if (o.precision != ORCTM_NH_FP64 && o.precision != ORCTM_NH_MIXED) return ORCTM_NH_ERR_INVALID;
  if (o.rtol_inner < 0) o.rtol_inner = (o.precision == ORCTM_NH_MIXED) ? 1e-5 : 0.0;
  if (o.maxit <= 0) o.maxit = 10000;
  if (o.max_outer <= 0) o.max_outer = 40;
  if (o.check_every <= 0) o.check_every = 16;
  if (!(o.rtol > 0)) o.rtol = 1e-10;
  if (o.tile_rows != 0 && o.tile_rows != 4 && o.tile_rows != 8 && o.tile_rows != 12) return ORCTM_NH_ERR_INVALID;
  if (o.bottom_solve != 0 && o.bottom_solve != 1) return ORCTM_NH_ERR_INVALID;
  if (o.zw_table < -1 || o.zw_table > 1) return ORCTM_NH_ERR_INVALID;
  if (o.buffer_addressing != 0 && o.buffer_addressing != -1) return ORCTM_NH_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= o.device || o.device < 0) return ORCTM_NH_ERR_NODEVICE;
  if (hipSetDevice(o.device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, o.device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ORCTM_NH_ERR_NODEVICE;
  orctm_nh_t* h = new (std::nothrow) orctm_nh_t();
  if (!h) return ORCTM_NH_ERR_NOMEM;
  h->opts = o;
  h->op = std::make_shared<OpData>();
  h->op->device = o.device;
  h->IE = grid->ie;
  h->JE = grid->je;
  h->KE = grid->ke;
  h->i_start = grid->i_start;
  h->j_start = grid->j_start;
  h->cyc_x = grid->cyclic_x ? 1 : 0;
  h->cyc_y = grid->cyclic_y ? 1 : 0;
  h->open_w = grid->open_w ? 1 : 0;
  h->open_e = grid->open_e ? 1 : 0;
  h->open_n = grid->open_n ? 1 : 0;
  h->open_s = grid->open_s ? 1 : 0;
  h->nx = grid->ie - 2;
  h->ny = grid->je - 2;
  h->nz = grid->ke;
  h->pitch = (h->nx + 31) / 32 * 32;
  h->plane = (long long)h->pitch * h->ny;
  h->fp32 = (o.precision == ORCTM_NH_MIXED);
  int rc = map_err(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (rc == ORCTM_NH_OK) rc = create_device(h, grid);
  if (rc == ORCTM_NH_OK && o.bottom_solve) rc = map_err(bot_create(h, grid));
  if (rc != ORCTM_NH_OK) {
    release(h);
    return rc;
  }
  *handle = h;
  return ORCTM_NH_OK;
}

int orctm_nh_destroy(orctm_nh_t* h) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  DEVLOCK;
  hipSetDevice(h->opts.device);
  release(h);
  return ORCTM_NH_OK;
}

int orctm_nh_set_rhs(orctm_nh_t* h, const double* divg) {
  if (!h || !divg) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  const size_t n = (size_t)h->IE * h->JE * h->KE;
  const int rc = ensure_ctx(h);
  if (rc != ORCTM_NH_OK) return rc;
  hipError_t e = ensure_fort(h);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_fort, divg, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = launch_rhs(h->stream, h->ctx.d_sys, h->ctx.d_tiles, h->ctx.ntiles, 1, h->d_fort, h->IE, h->JE);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return map_err(e);
}

int orctm_nh_set_guess(orctm_nh_t* h, const double* pnh) {
  if (!h || !pnh) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  const size_t n = (size_t)h->IE * h->JE * h->KE;
  const int rc = ensure_ctx(h);
  if (rc != ORCTM_NH_OK) return rc;
  hipError_t e = ensure_fort(h);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_fort, pnh, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = launch_guess(h->stream, h->ctx.d_sys, h->ctx.d_tiles, h->ctx.ntiles, h->d_fort, h->IE, h->JE);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) h->guess_set = true;
  return map_err(e);
}

int orctm_nh_solve_resident(orctm_nh_t* h, orctm_nh_stats_t* st) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  const double t0 = now_ms();
  hipSetDevice(h->opts.device);
  int rc = ensure_ctx(h);
  if (rc == ORCTM_NH_OK) rc = prepare_guess(h);
  if (rc != ORCTM_NH_OK) return rc;
  float ms = 0.f;
  hipError_t e = ctx_solve(h->ctx, &ms);
  if (e != hipSuccess) return map_err(e);
  fill_stats(h->ctx.h_outer[0], h->ctx.heights[0], h->ctx.tab[0], h->ctx.buf, st, ms, now_ms() - t0);
  return outer_status(h->ctx.h_outer[0]);
}

int orctm_nh_get_solution(orctm_nh_t* h, double* pnh) {
  if (!h || !pnh) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  // interior only (mo_petsc.f90:223-227); the halo of pnh is left as it is
  for (int k = 0; k < h->nz; ++k) {
    double* dst = pnh + 1 + (size_t)h->IE * (1 + (size_t)h->JE * k);
    const double* src = h->sys.p64 + (size_t)k * h->plane;
    hipError_t e = hipMemcpy2D(dst, (size_t)h->IE * 8, src, (size_t)h->pitch * 8, (size_t)h->nx * 8, (size_t)h->ny,
                               hipMemcpyDeviceToHost);
    if (e != hipSuccess) return map_err(e);
  }
  return ORCTM_NH_OK;
}

int orctm_nh_solve(orctm_nh_t* h, const double* divg, double* pnh, orctm_nh_stats_t* st) {
  if (!h || !divg || !pnh) return ORCTM_NH_ERR_INVALID;
  const double t0 = now_ms();
  int rc = orctm_nh_set_rhs(h, divg);
  if (rc != ORCTM_NH_OK) return rc;
  orctm_nh_stats_t s;
  const int status = orctm_nh_solve_resident(h, &s);
  if (status == ORCTM_NH_ERR_HIP || status == ORCTM_NH_ERR_NOMEM || status == ORCTM_NH_ERR_INVALID) return status;
  if (status != ORCTM_NH_ERR_NONFINITE) {
    rc = orctm_nh_get_solution(h, pnh);
    if (rc != ORCTM_NH_OK) return rc;
  }
  s.total_ms = now_ms() - t0;
  if (st) *st = s;
  return status;
}

int orctm_nh_apply(orctm_nh_t* h, const double* x, double* y) {
  if (!h || !x || !y) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  const size_t bytes = (size_t)h->plane * h->nz * sizeof(double);
  double *dx = nullptr, *dy = nullptr;
  hipError_t e = hipMalloc(&dx, bytes);
  if (e == hipSuccess) e = hipMalloc(&dy, bytes);
  if (e == hipSuccess) e = hipMemset(dx, 0, bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = to_device_vec(h, dx, x, sizeof(double));
  if (e == hipSuccess) e = launch_apply64(h->stream, h->ctx.d_sys, h->ctx.d_tiles, h->ctx.ntiles, dx, dy);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = to_host_vec(h, y, dy, sizeof(double));
  hipFree(dx);
  hipFree(dy);
  return map_err(e);
}

int orctm_nh_apply_noalloc(orctm_nh_t* h, const double* x, double* y) {
  if (!h || !x || !y) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  const size_t cells = (size_t)h->plane * h->nz;  // padded vector length, as orctm_nh_apply stages it
  if (cells > kScratchCells) {
    snprintf(g_last_error, sizeof(g_last_error),
             "orctm_nh_apply_noalloc: %zu padded cells (pitch %d x %d x %d) exceed the staging limit of %zu",
             cells, h->pitch, h->ny, h->nz, kScratchCells);
    return ORCTM_NH_ERR_INVALID;
  }
  if (hipSetDevice(h->opts.device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  double* base = nullptr;
  hipError_t e = scratch_address(h->opts.device, &base);
  if (e != hipSuccess) return map_err(e);
  // k_apply64 reads x and writes y at interior cells only, and only interior cells are copied in and
  // out, so nothing in the staging needs clearing between calls.
  double* dx = base;
  double* dy = base + kScratchCells;
  hipStream_t st = h->stream;
  const size_t es = sizeof(double);
  e = hipMemcpy2DAsync(dx, (size_t)h->pitch * es, x, (size_t)h->nx * es, (size_t)h->nx * es, (size_t)h->ny * h->nz,
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_apply64(st, h->ctx.d_sys, h->ctx.d_tiles, h->ctx.ntiles, dx, dy);
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(y, (size_t)h->nx * es, dy, (size_t)h->pitch * es, (size_t)h->nx * es, (size_t)h->ny * h->nz,
                         hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return map_err(e);
}

int orctm_nh_get_operator(orctm_nh_t* h, double* coef7) {
  if (!h || !coef7) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  const size_t NP = (size_t)h->plane * h->nz, N = (size_t)h->nx * h->ny * h->nz;
  double* dA = nullptr;
  hipError_t e = hipMalloc(&dA, 7 * NP * sizeof(double));
  if (e == hipSuccess) e = hipMemset(dA, 0, 7 * NP * sizeof(double));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = launch_materialize(h->stream, h->ctx.d_sys, h->ctx.d_tiles, h->ctx.ntiles, dA);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  for (int q = 0; q < 7 && e == hipSuccess; ++q) e = to_host_vec(h, coef7 + q * N, dA + q * NP, sizeof(double));
  hipFree(dA);
  return map_err(e);
}

int orctm_nh_get_dims(orctm_nh_t* h, int32_t* nx, int32_t* ny, int32_t* nz) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  if (nx) *nx = h->nx;
  if (ny) *ny = h->ny;
  if (nz) *nz = h->nz;
  return ORCTM_NH_OK;
}

int orctm_nh_device_bytes(orctm_nh_t* h, int64_t* bytes) {
  if (!h || !bytes) return ORCTM_NH_ERR_INVALID;
  *bytes = (int64_t)(h->bytes + (h->op ? h->op->bytes : 0));
  return ORCTM_NH_OK;
}

int orctm_nh_inner_apply(orctm_nh_t* h, const void* x, void* y) {
  if (!h || !x || !y) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  const size_t es = h->fp32 ? 4 : 8;
  Ctx& c = h->ctx;
  // x goes where the forward elimination leaves w; the matvec kernel in raw mode takes it as z
  hipError_t e = hipMemset(h->sys.z, 0, (size_t)h->plane * h->nz * es);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = to_device_vec(h, h->sys.z, x, es);
  if (e == hipSuccess) {
    if (h->fp32) {
      Inner<float> in{};
      in.first = 1;
      in.done = RUN;
      in.rho = 1.f;
      e = hipMemcpy(c.d_inner, &in, sizeof(in), hipMemcpyHostToDevice);
    } else {
      Inner<double> in{};
      in.first = 1;
      in.done = RUN;
      in.rho = 1.0;
      e = hipMemcpy(c.d_inner, &in, sizeof(in), hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess) e = iter_bwd(c, 1);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess) e = to_host_vec(h, y, h->sys.q, es);
  return map_err(e);
}

int orctm_nh_inner_precond(orctm_nh_t* h, const void* r, void* z, double* rho) {
  if (!h || !r || !z) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  const size_t es = h->fp32 ? 4 : 8;
  Ctx& c = h->ctx;
  // forward elimination (init kernel) then back substitution inside the matvec kernel, whose first
  // iteration writes p = z into search-direction buffer B
  Outer o{};
  o.odone = RUN;
  hipError_t e = hipMemcpy(c.d_outer, &o, sizeof(o), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = to_device_vec(h, h->sys.r, r, es);
  if (e == hipSuccess) e = iter_fwd(c, true);
  if (e == hipSuccess) e = iter_bwd(c, 0);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess) e = to_host_vec(h, z, h->sys.pb, es);
  if (e == hipSuccess && rho) {
    if (h->fp32) {
      Inner<float> in;
      e = hipMemcpy(&in, c.d_inner, sizeof(in), hipMemcpyDeviceToHost);
      *rho = in.rho;
    } else {
      Inner<double> in;
      e = hipMemcpy(&in, c.d_inner, sizeof(in), hipMemcpyDeviceToHost);
      *rho = in.rho;
    }
  }
  return map_err(e);
}

int orctm_nh_state_bytes(orctm_nh_t* h, int64_t* bytes) {
  if (!h || !bytes) return ORCTM_NH_ERR_INVALID;
  *bytes = (int64_t)h->bytes;
  return ORCTM_NH_OK;
}

int orctm_nh_clone(orctm_nh_t* src, orctm_nh_t** out) {
  if (!src || !out) return ORCTM_NH_ERR_INVALID;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(src->mu);
  DEVLOCK;
  hipSetDevice(src->opts.device);
  orctm_nh_t* h = new (std::nothrow) orctm_nh_t();
  if (!h) return ORCTM_NH_ERR_NOMEM;
  h->opts = src->opts;
  h->op = src->op;
  h->IE = src->IE;
  h->JE = src->JE;
  h->KE = src->KE;
  h->nx = src->nx;
  h->ny = src->ny;
  h->nz = src->nz;
  h->pitch = src->pitch;
  h->plane = src->plane;
  h->fp32 = src->fp32;
  h->i_start = src->i_start;
  h->j_start = src->j_start;
  h->cyc_x = src->cyc_x;
  h->cyc_y = src->cyc_y;
  h->open_w = src->open_w;
  h->open_e = src->open_e;
  h->open_n = src->open_n;
  h->open_s = src->open_s;
  h->sys = src->sys;  // operator pointers shared; state pointers replaced below
  h->d_smax = src->d_smax;
  int rc = map_err(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (rc == ORCTM_NH_OK) rc = alloc_state(h);
  if (rc == ORCTM_NH_OK) {
    // src's right-hand side was completed on src's stream before its entry point returned
    const size_t NP = (size_t)h->plane * h->nz;
    hipError_t e = hipMemcpyAsync(h->sys.b, src->sys.b, NP * sizeof(double), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->sys.nbp, src->sys.nbp, sizeof(double), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    rc = map_err(e);
  }
  if (rc != ORCTM_NH_OK) {
    release(h);
    return rc;
  }
  *out = h;
  return ORCTM_NH_OK;
}

int orctm_nh_perturb_rhs(orctm_nh_t* h, double amplitude, uint64_t seed) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  return map_err(ctx_perturb(h->ctx, amplitude, &seed));
}

int orctm_nh_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps) {
  if (bytes < 16 || reps <= 0 || !gbps) return ORCTM_NH_ERR_INVALID;
  DEVLOCK;
  if (hipSetDevice(device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  hipError_t e = hipMalloc(&a, (size_t)bytes);
  if (e == hipSuccess) e = hipMalloc(&b, (size_t)bytes);
  if (e == hipSuccess) e = hipMemset(a, 0, (size_t)bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_copy4(st, a, b, (size_t)bytes);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy4(st, a, b, (size_t)bytes);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *gbps = 2.0 * (double)(bytes / 16 * 16) * reps / (ms * 1e-3) / 1e9;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  hipFree(a);
  hipFree(b);
  return map_err(e);
}

int orctm_nh_attach_vgrid(orctm_nh_t* h, const orctm_nh_vgrid_t* vg) {
  if (!h || !vg) return ORCTM_NH_ERR_INVALID;
  const double* in2[] = {vg->dlxp, vg->dlyp, vg->ddpo, vg->weto, vg->amsuo, vg->dduo, vg->dtdxuo,
                         vg->dlyu, vg->amsue, vg->ddue, vg->dpye, vg->dlxv, vg->dtdxpo, vg->dtdyo};
  for (const double* q : in2)
    if (!q) return ORCTM_NH_ERR_INVALID;
  if (!(vg->dt > 0.0) || !std::isfinite(vg->dti)) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  hipSetDevice(h->opts.device);
  if (h->ocv_ready) return ORCTM_NH_ERR_INVALID;  // attached once per handle
  const size_t IE = h->IE, JE = h->JE, KE = h->KE;
  const size_t NU = IE - h->i_start + 1, NV = JE - h->j_start + 1;
  const size_t p2 = IE * JE, u2 = NU * JE, v2 = IE * NV;
  OcvArrays& a = h->ocv;
  a = OcvArrays{};
  a.IE = (int)IE;
  a.JE = (int)JE;
  a.KE = (int)KE;
  a.i_start = h->i_start;
  a.j_start = h->j_start;
  a.cyc_x = h->cyc_x;
  a.cyc_y = h->cyc_y;
  a.open_w = h->open_w;
  a.open_e = h->open_e;
  a.open_n = h->open_n;
  a.open_s = h->open_s;
  a.dt = vg->dt;
  a.dti = vg->dti;
  hipError_t e = hipSuccess;
  auto put = [&](const double*& dst, const double* src, size_t n) {
    void* p = nullptr;
    if (e == hipSuccess) e = dalloc(h, &p, n * sizeof(double));
    if (e == hipSuccess) e = upload(p, src, n * sizeof(double));
    dst = static_cast<const double*>(p);
  };
  auto buf = [&](double*& dst, size_t n) {
    void* p = nullptr;
    if (e == hipSuccess) e = dalloc(h, &p, n * sizeof(double));
    dst = static_cast<double*>(p);
  };
  put(a.dlxp, vg->dlxp, p2);
  put(a.dlyp, vg->dlyp, p2);
  put(a.dtdxpo, vg->dtdxpo, p2);
  put(a.dtdyo, vg->dtdyo, p2);
  put(a.ddpo, vg->ddpo, p2 * KE);
  put(a.weto, vg->weto, p2 * KE);
  put(a.amsuo, vg->amsuo, u2 * KE);
  put(a.dduo, vg->dduo, u2 * KE);
  put(a.dtdxuo, vg->dtdxuo, u2);
  put(a.dlyu, vg->dlyu, u2);
  put(a.amsue, vg->amsue, v2 * KE);
  put(a.ddue, vg->ddue, v2 * KE);
  put(a.dpye, vg->dpye, v2);
  put(a.dlxv, vg->dlxv, v2);
  buf(h->d_uoo, u2 * KE);
  buf(h->d_uoo_1, u2 * KE);
  buf(h->d_voe, v2 * KE);
  buf(h->d_voe_1, v2 * KE);
  buf(h->d_wno, p2 * (KE + 1));
  buf(h->d_wno_1, p2 * (KE + 1));
  buf(h->d_zo, p2);
  if (e == hipSuccess) e = ensure_fort(h);
  if (e != hipSuccess) return map_err(note_error(e, "attach_vgrid", __LINE__));
  a.uoo = h->d_uoo;
  a.uoo_1 = h->d_uoo_1;
  a.voe = h->d_voe;
  a.voe_1 = h->d_voe_1;
  a.wno = h->d_wno;
  a.wno_1 = h->d_wno_1;
  a.zo = h->d_zo;
  h->ocv_ready = true;
  return ORCTM_NH_OK;
}

int orctm_nh_divergence(orctm_nh_t* h, const double* uoo, const double* uoo_1, const double* voe,
                        const double* voe_1, const double* wno, const double* wno_1, const double* zo,
                        double* divg) {
  if (!h || !uoo || !uoo_1 || !voe || !voe_1 || !wno || !wno_1 || !zo) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  if (!h->ocv_ready) return ORCTM_NH_ERR_INVALID;
  hipSetDevice(h->opts.device);
  {
    const int crc = ensure_ctx(h);
    if (crc != ORCTM_NH_OK) return crc;
  }
  const size_t IE = h->IE, JE = h->JE, KE = h->KE;
  const size_t u3 = (IE - h->i_start + 1) * JE * KE, v3 = IE * (JE - h->j_start + 1) * KE, w3 = IE * JE * (KE + 1);
  hipStream_t st = h->stream;
  hipError_t e = hipMemcpyAsync(h->d_uoo, uoo, u3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_uoo_1, uoo_1, u3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_voe, voe, v3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_voe_1, voe_1, v3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_wno, wno, w3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_wno_1, wno_1, w3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_zo, zo, IE * JE * 8, hipMemcpyHostToDevice, st);
  // DIVG into the Fortran staging array; with the bottom solve, PNH_bot and the corrections of DIVG
  // (ComputeRHS with CG2D4NONHY); then b = DIVG*WETO exactly as orctm_nh_set_rhs forms it
  if (e == hipSuccess) e = launch_divergence(st, h->ocv, h->d_fort);
  int bot_status = ORCTM_NH_OK;
  if (e == hipSuccess && h->bot) {
    e = bot_step(st, *h->bot, h->ocv, h->d_fort, h->opts.rtol, h->opts.maxit);
    if (e == hipSuccess) e = bot_correct(st, *h->bot, h->ocv, h->d_fort);
    if (e == hipSuccess) bot_status = h->bot->status;
  }
  if (e == hipSuccess)
    e = launch_rhs(st, h->ctx.d_sys, h->ctx.d_tiles, h->ctx.ntiles, 1, h->d_fort, h->IE, h->JE);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && divg) {
    for (size_t k = 0; k < KE && e == hipSuccess; ++k) {
      const size_t off = 1 + IE * (1 + JE * k);
      e = hipMemcpy2D(divg + off, IE * 8, h->d_fort + off, IE * 8, (IE - 2) * 8, JE - 2, hipMemcpyDeviceToHost);
    }
  }
  return e == hipSuccess ? bot_status : map_err(e);
}

int orctm_nh_get_bottom(orctm_nh_t* h, double* pnh_bot, double* divg_bot_surf, int32_t* iterations,
                        double* rel_res, double* device_ms) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  if (!h->bot) return ORCTM_NH_ERR_INVALID;
  hipSetDevice(h->opts.device);
  const size_t IE = h->IE, JE = h->JE;
  hipError_t e = hipSuccess;
  if (pnh_bot) e = hipMemcpy(pnh_bot, h->bot->pnh_bot, IE * JE * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && divg_bot_surf)
    e = hipMemcpy2D(divg_bot_surf + 1 + IE, IE * 8, h->bot->dbs + 1 + IE, IE * 8, (IE - 2) * 8, JE - 2,
                    hipMemcpyDeviceToHost);
  if (iterations) *iterations = h->bot->iters < 0 ? 0 : h->bot->iters;
  if (rel_res) *rel_res = h->bot->rel_res;
  if (e == hipSuccess && device_ms) {
    float ms = 0.f;
    *device_ms = 0.0;
    if (h->bot->iters >= 0) e = hipEventElapsedTime(&ms, h->bot->ev0, h->bot->ev1);
    if (e == hipSuccess) *device_ms = ms;
  }
  return map_err(e);
}

int orctm_nh_bottom_symmetry(orctm_nh_t* h, double* defect) {
  if (!h || !defect) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  if (!h->bot) return ORCTM_NH_ERR_INVALID;
  *defect = h->bot->sym_defect;
  return ORCTM_NH_OK;
}

int orctm_nh_project(orctm_nh_t* h, double* pnh, double* uoo, double* voe, double* wno) {
  if (!h || !pnh || !uoo || !voe || !wno) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  if (!h->ocv_ready) return ORCTM_NH_ERR_INVALID;
  hipSetDevice(h->opts.device);
  const size_t IE = h->IE, JE = h->JE, KE = h->KE;
  const size_t p3 = IE * JE * KE, u3 = (IE - h->i_start + 1) * JE * KE, v3 = IE * (JE - h->j_start + 1) * KE,
               w3 = IE * JE * (KE + 1);
  hipStream_t st = h->stream;
  hipError_t e = hipMemcpyAsync(h->d_fort, pnh, p3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_uoo, uoo, u3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(h->d_voe, voe, v3 * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = launch_project(st, h->ocv, h->sys.p64, h->pitch, h->plane, h->d_fort, h->d_uoo, h->d_voe, h->d_wno);
  if (e == hipSuccess) e = hipMemcpyAsync(pnh, h->d_fort, p3 * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(uoo, h->d_uoo, u3 * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(voe, h->d_voe, v3 * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(wno, h->d_wno, w3 * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return map_err(e);
}

int orctm_nh_batch_create(int32_t n, orctm_nh_t* const* systems, orctm_nh_batch_t** batch) {
  if (!batch || n <= 0 || !systems) return ORCTM_NH_ERR_INVALID;
  DEVLOCK;
  *batch = nullptr;
  const bool fp32 = systems[0] ? systems[0]->fp32 : true;
  int max_outer = 0, check = 16, dev = 0;
  bool prof = false;
  std::vector<Sys> ss;
  std::vector<int> forced, zwt, adr;
  for (int q = 0; q < n; ++q) {
    orctm_nh_t* h = systems[q];
    if (!h || h->fp32 != fp32 || h->opts.device != systems[0]->opts.device) return ORCTM_NH_ERR_INVALID;
    for (int r = 0; r < q; ++r)
      if (systems[r] == h) return ORCTM_NH_ERR_INVALID;
    ss.push_back(h->sys);
    forced.push_back(h->opts.tile_rows);
    zwt.push_back(h->opts.zw_table);
    adr.push_back(h->opts.buffer_addressing);
    max_outer = std::max(max_outer, h->opts.max_outer);
    check = systems[0]->opts.check_every;
    dev = h->opts.device;
    prof |= h->opts.profile != 0;
  }
  orctm_nh_batch_t* b = new (std::nothrow) orctm_nh_batch_t();
  if (!b) return ORCTM_NH_ERR_NOMEM;
  b->members.assign(systems, systems + n);
  hipSetDevice(dev);
  hipError_t e = ctx_build(b->ctx, ss, forced, zwt, adr, fp32, check, max_outer, dev, prof, nullptr);
  if (e != hipSuccess) {
    if (b->ctx.stream) hipStreamSynchronize(b->ctx.stream);
    ctx_destroy(b->ctx);
    delete b;
    return map_err(e);
  }
  *batch = b;
  return ORCTM_NH_OK;
}

int orctm_nh_batch_solve(orctm_nh_batch_t* b, orctm_nh_stats_t* st) {
  if (!b) return ORCTM_NH_ERR_INVALID;
  const double t0 = now_ms();
  hipSetDevice(b->ctx.device);
  for (orctm_nh_t* h : b->members) {
    if (!h->opts.warm_start && !h->guess_set) {
      hipError_t e = hipMemsetAsync(h->sys.p64, 0, (size_t)h->plane * h->nz * sizeof(double), b->ctx.stream);
      if (e != hipSuccess) return map_err(e);
    }
    h->guess_set = false;
  }
  float ms = 0.f;
  hipError_t e = ctx_solve(b->ctx, &ms);
  if (e != hipSuccess) return map_err(e);
  const double tot = now_ms() - t0;
  int rc = ORCTM_NH_OK;
  for (int q = 0; q < b->ctx.nsys; ++q) {
    const int s = outer_status(b->ctx.h_outer[q]);
    if (rc == ORCTM_NH_OK && s != ORCTM_NH_OK) rc = s;
    if (st) fill_stats(b->ctx.h_outer[q], b->ctx.heights[q], b->ctx.tab[q], b->ctx.buf, st + q, ms, tot);
  }
  return rc;
}

}  // extern "C"

int onh::handle_precision(const orctm_nh_t* h) { return h ? (h->fp32 ? ORCTM_NH_MIXED : ORCTM_NH_FP64) : -1; }

void onh::batch_use_graphs(orctm_nh_batch_t* b, bool on) {
  if (b) b->ctx.use_graphs = on;
}

int onh::batch_perturb(orctm_nh_batch_t* b, double amplitude, const uint64_t* seeds) {
  if (!b || !seeds) return ORCTM_NH_ERR_INVALID;
  hipSetDevice(b->ctx.device);
  return map_err(ctx_perturb(b->ctx, amplitude, seeds));
}

extern "C" {

static int profile_read(Ctx& c, orctm_nh_profile_t* p, int32_t reset) {
  if (!p) return ORCTM_NH_ERR_INVALID;
  memset(p, 0, sizeof(*p));
  p->matvec_ms = c.mv_ms;
  p->update_ms = c.up_ms;
  p->matvec_launches = c.mv_n;
  p->update_launches = c.up_n;
  p->cells = c.cells;
  if (reset) {
    c.mv_ms = c.up_ms = 0.0;
    c.mv_n = c.up_n = 0;
  }
  return c.prof ? ORCTM_NH_OK : ORCTM_NH_ERR_INVALID;
}

int orctm_nh_profile_read(orctm_nh_t* h, orctm_nh_profile_t* p, int32_t reset) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  return profile_read(h->ctx, p, reset);
}

int orctm_nh_set_profile(orctm_nh_t* h, int32_t on) {
  if (!h) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(h->mu);
  DEVLOCK;
  const int want = on ? 1 : 0;
  if ((h->opts.profile != 0) == (want != 0)) return ORCTM_NH_OK;
  hipSetDevice(h->opts.device);
  if (h->stream) hipStreamSynchronize(h->stream);
  ctx_destroy(h->ctx);
  h->ctx_ready = false;
  h->opts.profile = want;
  return ensure_ctx(h);
}

int orctm_nh_batch_profile_read(orctm_nh_batch_t* b, orctm_nh_profile_t* p, int32_t reset) {
  if (!b) return ORCTM_NH_ERR_INVALID;
  return profile_read(b->ctx, p, reset);
}

int orctm_nh_batch_destroy(orctm_nh_batch_t* b) {
  if (!b) return ORCTM_NH_ERR_INVALID;
  DEVLOCK;
  hipSetDevice(b->ctx.device);
  if (b->ctx.stream) hipStreamSynchronize(b->ctx.stream);
  ctx_destroy(b->ctx);
  delete b;
  return ORCTM_NH_OK;
}

}  // extern "C"
