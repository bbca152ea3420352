// nh_serve.cpp -- continuous-batching request loop over the solver's C ABI.
//
// A request = one grid (a "model" handle whose assembled operator and resident right-hand side
// define it) and a number of consecutive time steps. Between steps the right-hand side is perturbed
// on the device and the solve warm-starts from the previous pressure, which stays resident, as
// KSPSetInitialGuessNonzero does between model time steps (mo_petsc.f90:232).
//
// The server thread admits queued requests while the device-memory budget and the slot count allow
// (a request's state is a clone of its model: operator shared, ~40 B/cell of its own), then
// advances every admitted request by one step in one batched solve (one launch per solver phase for
// all of them, per-request convergence masks), retires finished requests and admits new ones before
// the next round. Requests that arrive while a round runs join at the next round.
//
// Each request's arithmetic is independent of the others in the batch (per-system reductions in a
// fixed order, counter-based perturbation streams), so its result is bit-identical alone and in any
// mix. No counterpart in the reference: ORCTM runs one solve per model time step.
#include <hip/hip_runtime.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/orctm_nh.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t ticket = 0;
  orctm_nh_t* model = nullptr;
  orctm_nh_t* state = nullptr;  // clone, while admitted
  int32_t nsteps = 1;
  double perturb = 0.0;
  uint64_t seed = 0;
  double* pnh_out = nullptr;
  int64_t bytes = 0;
  orctm_nh_request_stats_t st{};
  bool done = false;
};

}  // namespace

struct orctm_nh_server {
  orctm_nh_server_options_t opts{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::vector<std::shared_ptr<Request>> active;
  std::map<int64_t, std::shared_ptr<Request>> all;
  int64_t next_ticket = 1;
  int64_t bytes_in_use = 0;
  int64_t rounds = 0;
  bool stop = false;
  std::thread worker;

  double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
  void run();
};

namespace {

void finish(orctm_nh_server_t* s, const std::shared_ptr<Request>& r, int status) {
  if (r->st.status == ORCTM_NH_OK && status != ORCTM_NH_OK) r->st.status = status;
  r->st.t_done_ms = s->now();
  r->done = true;
}

}  // namespace

void orctm_nh_server::run() {
  hipSetDevice(opts.device);
  for (;;) {
    std::vector<std::shared_ptr<Request>> admit_now;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
      if (stop && queue.empty() && active.empty()) return;
      // admission: FIFO, by slots and by device-memory budget
      while (!queue.empty() && (int)(active.size() + admit_now.size()) < opts.max_active) {
        std::shared_ptr<Request> r = queue.front();
        if (r->bytes > opts.hbm_budget) {  // can never fit: refuse, do not block the queue
          queue.pop_front();
          finish(this, r, ORCTM_NH_ERR_NOMEM);
          cv_done.notify_all();
          continue;
        }
        if (bytes_in_use + r->bytes > opts.hbm_budget) break;
        queue.pop_front();
        bytes_in_use += r->bytes;
        admit_now.push_back(r);
      }
    }
    for (auto& r : admit_now) {
      r->st.t_admit_ms = now();
      const int rc = orctm_nh_clone(r->model, &r->state);
      std::lock_guard<std::mutex> lk(mu);
      if (rc != ORCTM_NH_OK) {
        bytes_in_use -= r->bytes;
        finish(this, r, rc);
        cv_done.notify_all();
      } else {
        active.push_back(r);
      }
    }
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::lock_guard<std::mutex> lk(mu);
      batch = active;
    }
    if (batch.empty()) continue;
    // one step of every admitted request, in one batched solve
    std::vector<orctm_nh_t*> hs;
    for (auto& r : batch) hs.push_back(r->state);
    orctm_nh_batch_t* b = nullptr;
    int rc = orctm_nh_batch_create((int32_t)hs.size(), hs.data(), &b);
    std::vector<orctm_nh_stats_t> st(hs.size());
    if (rc == ORCTM_NH_OK) {
      orctm_nh_batch_solve(b, st.data());  // per-member statuses are in st
      orctm_nh_batch_destroy(b);
    } else {
      for (auto& x : st) x.status = rc;
    }
    ++rounds;
    std::vector<std::shared_ptr<Request>> retired;
    for (size_t q = 0; q < batch.size(); ++q) {
      Request& r = *batch[q];
      r.st.steps_done += 1;
      r.st.inner_iters += st[q].inner_iters;
      r.st.outer_iters += st[q].outer_iters;
      r.st.rel_res = st[q].rel_res;
      if (r.st.status == ORCTM_NH_OK && st[q].status != ORCTM_NH_OK) r.st.status = st[q].status;
      const bool last = r.st.steps_done >= r.nsteps || st[q].status >= ORCTM_NH_ERR_NONFINITE;
      if (last) {
        if (r.pnh_out) {
          const int g = orctm_nh_get_solution(r.state, r.pnh_out);
          if (g != ORCTM_NH_OK && r.st.status == ORCTM_NH_OK) r.st.status = g;
        }
        orctm_nh_destroy(r.state);
        r.state = nullptr;
        retired.push_back(batch[q]);
      } else {
        // next time step: perturbed right-hand side, warm start from the resident pressure
        const int p = orctm_nh_perturb_rhs(r.state, r.perturb, r.seed + (uint64_t)r.st.steps_done);
        if (p != ORCTM_NH_OK && r.st.status == ORCTM_NH_OK) r.st.status = p;
      }
    }
    if (!retired.empty()) {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& r : retired) {
        bytes_in_use -= r->bytes;
        finish(this, r, r->st.status);
        for (size_t q = 0; q < active.size(); ++q)
          if (active[q] == r) {
            active.erase(active.begin() + q);
            break;
          }
      }
      cv_done.notify_all();
    }
  }
}

extern "C" {

void orctm_nh_server_default_options(orctm_nh_server_options_t* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->hbm_budget = 0;
  o->max_active = 64;
  o->device = 0;
}

int orctm_nh_server_create(const orctm_nh_server_options_t* opts, orctm_nh_server_t** out) {
  if (!out) return ORCTM_NH_ERR_INVALID;
  *out = nullptr;
  orctm_nh_server_options_t o;
  orctm_nh_server_default_options(&o);
  if (opts) o = *opts;
  if (o.max_active <= 0) o.max_active = 64;
  if (hipSetDevice(o.device) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
  if (o.hbm_budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return ORCTM_NH_ERR_NODEVICE;
    o.hbm_budget = (int64_t)(0.9 * (double)fr);
  }
  orctm_nh_server_t* s = new orctm_nh_server();
  s->opts = o;
  s->t0 = Clock::now();
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return ORCTM_NH_OK;
}

int orctm_nh_server_submit(orctm_nh_server_t* s, orctm_nh_t* model, int32_t nsteps, double perturb, uint64_t seed,
                           double* pnh_out, int64_t* ticket) {
  if (!s || !model || nsteps <= 0 || !ticket) return ORCTM_NH_ERR_INVALID;
  int64_t bytes = 0;
  const int rc = orctm_nh_state_bytes(model, &bytes);
  if (rc != ORCTM_NH_OK) return rc;
  auto r = std::make_shared<Request>();
  r->model = model;
  r->nsteps = nsteps;
  r->perturb = perturb;
  r->seed = seed;
  r->pnh_out = pnh_out;
  r->bytes = bytes;
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
  std::shared_ptr<Request> r = it->second;
  s->cv_done.wait(lk, [&] { return r->done; });
  if (st) *st = r->st;
  return ORCTM_NH_OK;
}

int orctm_nh_server_drain(orctm_nh_server_t* s) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->active.empty(); });
  return ORCTM_NH_OK;
}

int orctm_nh_server_info(orctm_nh_server_t* s, int64_t* rounds, int64_t* bytes_in_use, int32_t* active,
                         int32_t* queued) {
  if (!s) return ORCTM_NH_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (rounds) *rounds = s->rounds;
  if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
  if (active) *active = (int32_t)s->active.size();
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
  if (s->worker.joinable()) s->worker.join();
  delete s;
  return ORCTM_NH_OK;
}

}  // extern "C"
