// Launcher of the single-workgroup coarse-level V-cycle kernel (mg_tail.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mg_plan.h"

namespace artss {

// Runs VCycleMG's descent and ascent over levels plan->first .. plan->last
// (device-resident plan) with `b_first` as the right-hand side of the first
// level; one workgroup, enqueued on `s`.
void mg_tail(hipStream_t s, const TailPlan *plan, const double *b_first);

}  // namespace artss
