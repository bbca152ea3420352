// Host harness for tests/test_serve_step.py: the serving loop's bookkeeping of one batched time step
// (orctm_amd/csrc/nh_serve_step.h), called on synthetic outcomes of orctm_nh_batch_solve.
#include "../orctm_amd/csrc/nh_serve_step.h"

extern "C" void serve_apply_step(int rc, const orctm_nh_stats_t* st, int n, const int32_t* nsteps,
                                 orctm_nh_request_stats_t* rs, int32_t* retire) {
  bool out[64] = {};
  onh::apply_step(rc, st, n, nsteps, rs, out);
  for (int q = 0; q < n; ++q) retire[q] = out[q] ? 1 : 0;
}
