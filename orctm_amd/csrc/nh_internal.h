// nh_internal.h -- library-internal entry points shared by the solver (nh_api.hip) and the serving
// loop (nh_serve.cpp); not part of the C ABI.
#pragma once

#include <stdint.h>

#include "../../include/orctm_nh.h"

namespace onh {

// Perturbs every member's resident right-hand side, member q with seeds[q] (the same stream of
// numbers as orctm_nh_perturb_rhs with that seed), in one launch over the batch; synchronous.
int batch_perturb(orctm_nh_batch_t* b, double amplitude, const uint64_t* seeds);

// Whether a batch replays its inner iterations from captured hipGraphs (the default) or launches them
// directly: a serving cohort lives for a few solves, where a capture (which also holds the process-wide
// device lock the other lanes' admissions need) costs more than the launches it saves.
void batch_use_graphs(orctm_nh_batch_t* b, bool on);

// The precision mode of a handle (ORCTM_NH_MIXED or ORCTM_NH_FP64); -1 for a null handle.
int handle_precision(const orctm_nh_t* h);

}  // namespace onh
