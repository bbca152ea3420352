// nh_serve.cpp -- continuous-batching request loop over the solver's C ABI.
//
// A request = one grid (a "model" handle whose assembled operator and resident right-hand side
// define it) and a number of consecutive time steps. Between steps the right-hand side is perturbed
// on the device and the solve warm-starts from the previous pressure, which stays resident, as
// KSPSetInitialGuessNonzero does between model time steps (mo_petsc.f90:232).
//
// Worker lanes, each a thread driving its own solver streams, run cohorts concurrently: a free lane
// admits the oldest queued request that fits the device-memory budget, with the queued requests of
// the same size class that also fit, clones their models (state from the library's memory pool, so
// admission and retirement never wait for the whole device), builds one batch over them and advances
// them one batched solve per time step (one launch per solver phase for the cohort, per-request
// convergence masks). A request retires after its last step; the other members carry on in a batch
// rebuilt over them. Lane 0 admits only small requests when there are several lanes, so a burst of
// large requests cannot hold every lane while small ones queue.
//
// Each request's arithmetic is independent of everything else on the device (per-system reductions
// in a fixed order, counter-based perturbation streams keyed by the request's own seed), so its result
// is bit-identical alone and in any mix. No counterpart in the reference: ORCTM runs one solve per
// model time step.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/orctm_nh.h"
#include "nh_internal.h"
#include "nh_pool.h"
#include "nh_serve_step.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr long long kSmallCells = 1LL << 24;   // lane 0 takes only requests below this size
constexpr long long kCohortCells = 1LL << 25;  // a cohort grows (same size class) up to this many cells
constexpr int kMaxLanes = 16;

struct Request {
  int64_t ticket = 0;
  orctm_nh_t* model = nullptr;
  orctm_nh_t* state = nullptr;  // clone, while admitted
  int32_t nsteps = 1;
  double perturb = 0.0;
  uint64_t seed = 0;
  double* pnh_out = nullptr;
  int64_t bytes = 0;
  long long cells = 0;
  int cls = 0;     // floor(log2(cells))
  int32_t prec = 0;
  orctm_nh_request_stats_t st{};
  bool done = false;
};

using ReqP = std::shared_ptr<Request>;

int floor_log2(long long v) {
  int c = 0;
  while ((2LL << c) <= v) ++c;
  return c;
}

}  // namespace

struct orctm_nh_server {
  orctm_nh_server_options_t opts{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ReqP> queue;
  std::map<int64_t, ReqP> all;
  int64_t next_ticket = 1;
  int64_t bytes_in_use = 0;
  int64_t rounds = 0;
  int active = 0;
  bool stop = false;
  std::vector<std::thread> lanes;

  double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
  void lane(int id);
  bool small_only(int id) const { return id == 0 && opts.lanes > 1; }
  // Under mu: index of the request this lane would admit next (-1: none can be admitted now).
  // Requests that can never fit are refused on the way.
  int pick(int id);
  void finish(const ReqP& r, int status);  // under mu
  void run_cohort(std::vector<ReqP>& cohort);
};

void orctm_nh_server::finish(const ReqP& r, int status) {
  if (r->st.status == ORCTM_NH_OK && status != ORCTM_NH_OK) r->st.status = status;
  r->st.t_done_ms = now();
  r->done = true;
}

int orctm_nh_server::pick(int id) {
  for (size_t q = 0; q < queue.size();) {
    const ReqP& r = queue[q];
    if (r->bytes > opts.hbm_budget) {  // can never fit: refuse it, do not block the queue
      ReqP x = r;
      queue.erase(queue.begin() + q);
      finish(x, ORCTM_NH_ERR_NOMEM);
      cv_done.notify_all();
      continue;
    }
    if (small_only(id) && r->cells >= kSmallCells) {
      ++q;
      continue;
    }
    // FIFO: the first request this lane may take is admitted when it fits, else the lane waits
    if (active >= opts.max_active || bytes_in_use + r->bytes > opts.hbm_budget) return -1;
    return (int)q;
  }
  return -1;
}

void orctm_nh_server::lane(int id) {
  hipSetDevice(opts.device);
  for (;;) {
    std::vector<ReqP> cohort;
    {
      std::unique_lock<std::mutex> lk(mu);
      int q = -1;
      cv_work.wait(lk, [&] {
        q = pick(id);
        return q >= 0 || (stop && queue.empty());
      });
      if (q < 0) return;  // stopping and nothing left to admit
      ReqP head = queue[q];
      queue.erase(queue.begin() + q);
      cohort.push_back(head);
      bytes_in_use += head->bytes;
      ++active;
      long long cells = head->cells;
      // the queued requests of the same size class and precision that fit join the cohort
      for (size_t k = 0; k < queue.size() && (int)cohort.size() < opts.max_cohort;) {
        const ReqP& r = queue[k];
        if (r->cls == head->cls && r->prec == head->prec && cells + r->cells <= kCohortCells &&
            active < opts.max_active && bytes_in_use + r->bytes <= opts.hbm_budget) {
          cohort.push_back(r);
          bytes_in_use += r->bytes;
          ++active;
          cells += r->cells;
          queue.erase(queue.begin() + k);
        } else {
          ++k;
        }
      }
    }
    run_cohort(cohort);
    cv_work.notify_all();  // memory and slots were released: other lanes may admit now
  }
}

void orctm_nh_server::run_cohort(std::vector<ReqP>& cohort) {
  std::vector<ReqP> members;
  for (ReqP& r : cohort) {
    r->st.t_admit_ms = now();
    const int rc = orctm_nh_clone(r->model, &r->state);
    if (rc != ORCTM_NH_OK) {
      std::lock_guard<std::mutex> lk(mu);
      bytes_in_use -= r->bytes;
      --active;
      finish(r, rc);
      cv_done.notify_all();
    } else {
      members.push_back(r);
    }
  }
  bool perturb_first = false;  // members whose next step needs its perturbation (after a rebuild)
  while (!members.empty()) {
    std::vector<orctm_nh_t*> hs;
    for (ReqP& r : members) hs.push_back(r->state);
    orctm_nh_batch_t* b = nullptr;
    int rc = orctm_nh_batch_create((int32_t)hs.size(), hs.data(), &b);
    if (rc == ORCTM_NH_OK) onh::batch_use_graphs(b, opts.cohort_graphs != 0);
    std::vector<uint64_t> seeds(members.size());
    auto perturb_all = [&]() {
      for (size_t q = 0; q < members.size(); ++q) seeds[q] = members[q]->seed + (uint64_t)members[q]->st.steps_done;
      // one amplitude per launch: members with the same amplitude share it (the common case)
      const double amp = members[0]->perturb;
      bool same = true;
      for (ReqP& r : members) same &= (r->perturb == amp);
      if (same) return onh::batch_perturb(b, amp, seeds.data());
      for (size_t q = 0; q < members.size(); ++q) {
        const int p = orctm_nh_perturb_rhs(members[q]->state, members[q]->perturb, seeds[q]);
        if (p != ORCTM_NH_OK) return p;
      }
      return (int)ORCTM_NH_OK;
    };
    if (rc == ORCTM_NH_OK && perturb_first) rc = perturb_all();
    std::vector<ReqP> retired;
    while (rc == ORCTM_NH_OK) {
      const int n = (int)members.size();
      std::vector<orctm_nh_stats_t> st(n);
      const int s = orctm_nh_batch_solve(b, st.data());
      std::vector<orctm_nh_request_stats_t> rs(n);
      std::vector<int32_t> nsteps(n);
      std::unique_ptr<bool[]> out(new bool[n]());
      for (int q = 0; q < n; ++q) {
        rs[q] = members[q]->st;
        nsteps[q] = members[q]->nsteps;
      }
      onh::apply_step(s, st.data(), n, nsteps.data(), rs.data(), out.get());
      {
        std::lock_guard<std::mutex> lk(mu);
        ++rounds;
        for (int q = 0; q < n; ++q) members[q]->st = rs[q];
      }
      for (int q = 0; q < n; ++q)
        if (out[q]) retired.push_back(members[q]);
      if (onh::batch_error(s)) break;  // every member retires with the batch's error
      if (!retired.empty()) break;
      rc = perturb_all();  // next time step of every member: perturbed right-hand side, warm start
    }
    if (b) orctm_nh_batch_destroy(b);
    if (rc != ORCTM_NH_OK) retired = members;  // the cohort cannot go on: every member retires with rc
    for (ReqP& r : retired) {
      int status = rc;
      if (r->pnh_out && r->st.steps_done > 0) {
        const int g = orctm_nh_get_solution(r->state, r->pnh_out);
        if (status == ORCTM_NH_OK) status = g;
      }
      orctm_nh_destroy(r->state);
      r->state = nullptr;
      std::lock_guard<std::mutex> lk(mu);
      bytes_in_use -= r->bytes;
      --active;
      finish(r, status);  // keeps a status the request already carries
      members.erase(std::find(members.begin(), members.end(), r));
    }
    if (!retired.empty()) {
      cv_done.notify_all();
      cv_work.notify_all();
    }
    perturb_first = true;
  }
}

extern "C" {

void orctm_nh_server_default_options(orctm_nh_server_options_t* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->hbm_budget = 0;
  o->max_active = 64;
  o->device = 0;
  o->lanes = 4;
  o->max_cohort = 16;
}

int orctm_nh_server_create(const orctm_nh_server_options_t* opts, orctm_nh_server_t** out) {
  if (!out) return ORCTM_NH_ERR_INVALID;
  *out = nullptr;
  orctm_nh_server_options_t o;
  orctm_nh_server_default_options(&o);
  if (opts) o = *opts;
  if (o.max_active <= 0) o.max_active = 64;
  if (o.lanes <= 0) o.lanes = 4;
  if (o.max_cohort <= 0) o.max_cohort = 16;
  if (o.lanes > kMaxLanes) return ORCTM_NH_ERR_INVALID;
  if (hipSetDevice(o.device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  if (o.hbm_budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
    o.hbm_budget = (int64_t)(0.9 * (double)(fr + onh::pool_idle_bytes(o.device)));
  }
  orctm_nh_server_t* s = new (std::nothrow) orctm_nh_server();
  if (!s) return ORCTM_NH_ERR_NOMEM;
  s->opts = o;
  s->t0 = Clock::now();
  for (int k = 0; k < o.lanes; ++k) s->lanes.emplace_back([s, k] { s->lane(k); });
  *out = s;
  return ORCTM_NH_OK;
}

int orctm_nh_server_submit(orctm_nh_server_t* s, orctm_nh_t* model, int32_t nsteps, double perturb, uint64_t seed,
                           double* pnh_out, int64_t* ticket) {
  if (!s || !model || nsteps <= 0 || !ticket) return ORCTM_NH_ERR_INVALID;
  int64_t bytes = 0;
  int rc = orctm_nh_state_bytes(model, &bytes);
  if (rc != ORCTM_NH_OK) return rc;
  int32_t nx = 0, ny = 0, nz = 0;
  rc = orctm_nh_get_dims(model, &nx, &ny, &nz);
  if (rc != ORCTM_NH_OK) return rc;
  auto r = std::make_shared<Request>();
  r->model = model;
  r->nsteps = nsteps;
  r->perturb = perturb;
  r->seed = seed;
  r->pnh_out = pnh_out;
  r->bytes = bytes;
  r->cells = (long long)nx * ny * nz;
  r->cls = floor_log2(std::max(r->cells, 1LL));
  r->prec = onh::handle_precision(model);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return ORCTM_NH_ERR_INVALID;
  r->ticket = s->next_ticket++;
  r->st.t_submit_ms = s->now();
  s->queue.push_back(r);
  s->all[r->ticket] = r;
  *ticket = r->ticket;
  s->cv_work.notify_all();
  return ORCTM_NH_OK;
}

int orctm_nh_server_poll(orctm_nh_server_t* s, int64_t ticket, orctm_nh_request_stats_t* st) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return ORCTM_NH_ERR_INVALID;
  if (st) *st = it->second->st;
  return it->second->done ? ORCTM_NH_OK : ORCTM_NH_PENDING;
}

int orctm_nh_server_wait(orctm_nh_server_t* s, int64_t ticket, orctm_nh_request_stats_t* st) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return ORCTM_NH_ERR_INVALID;
  ReqP r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  if (st) *st = r->st;
  return ORCTM_NH_OK;
}

int orctm_nh_server_drain(orctm_nh_server_t* s) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->active == 0; });
  return ORCTM_NH_OK;
}

int orctm_nh_server_info(orctm_nh_server_t* s, int64_t* rounds, int64_t* bytes_in_use, int32_t* active,
                         int32_t* queued) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (rounds) *rounds = s->rounds;
  if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
  if (active) *active = (int32_t)s->active;
  if (queued) *queued = (int32_t)s->queue.size();
  return ORCTM_NH_OK;
}

int orctm_nh_server_destroy(orctm_nh_server_t* s) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->lanes)
    if (t.joinable()) t.join();
  delete s;
  return ORCTM_NH_OK;
}

}  // extern "C"
