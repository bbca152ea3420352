// nh_serve_step.h -- the serving loop's bookkeeping of one batched solve (host code only).
//
// Kept apart from the loop so that the CPU test suite can compile and check it without a device.
#pragma once

#include <stdint.h>

#include "../../include/orctm_nh.h"

namespace onh {

// A batch-wide failure of orctm_nh_batch_solve (the solve could not run, no member stats were written)
// rather than the first non-OK status of a member.
inline bool batch_error(int rc) {
  return rc == ORCTM_NH_ERR_INVALID || rc == ORCTM_NH_ERR_NOMEM || rc == ORCTM_NH_ERR_HIP ||
         rc == ORCTM_NH_ERR_NODEVICE;
}

// Applies the outcome of one batched time step of n cohort members to their request statistics:
// rc is what orctm_nh_batch_solve returned, st its per-member statistics, nsteps each request's number
// of time steps. retire[q] is set for the members that leave the cohort after this step: the request's
// last step, a fatal member status (non-finite right-hand side or an error), or a batch-wide failure,
// in which case no step was done and every member carries rc as its status.
inline void apply_step(int rc, const orctm_nh_stats_t* st, int n, const int32_t* nsteps,
                       orctm_nh_request_stats_t* rs, bool* retire) {
  for (int q = 0; q < n; ++q) {
    if (batch_error(rc)) {
      if (rs[q].status == ORCTM_NH_OK) rs[q].status = rc;
      retire[q] = true;
      continue;
    }
    rs[q].steps_done += 1;
    rs[q].inner_iters += st[q].inner_iters;
    rs[q].outer_iters += st[q].outer_iters;
    rs[q].rel_res = st[q].rel_res;
    if (rs[q].status == ORCTM_NH_OK && st[q].status != ORCTM_NH_OK) rs[q].status = st[q].status;
    retire[q] = rs[q].steps_done >= nsteps[q] || st[q].status >= ORCTM_NH_ERR_NONFINITE;
  }
}

}  // namespace onh
