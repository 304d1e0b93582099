// Launcher of the single-workgroup coarse-level V-cycle kernel (mg_tail.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mg_plan.h"

namespace artss {

// Runs VCycleMG's descent and ascent over levels plan->first .. plan->last
// (device-resident plan) with `b_first` as the right-hand side of the first
// level; one workgroup, enqueued on `s`.
// max_iter, n_relax, w > 0 replace the plan's coarse-solve cap, smoothing
// count and Jacobi weight for this launch; skip_dead leaves out the
// reference's work on the coarsest level, whose error stays zero (the
// result differs at most in the sign of zeros); coarse_tol sets the coarse
// solve's stopping test: 0 the plan's absolute tolerance (the reference's),
// > 0 that factor times the residual of the solve's starting iterate, < 0
// no test (max_iter sweeps)
void mg_tail(hipStream_t s, const TailPlan *plan, const double *b_first, int lds_bytes, int max_iter = 0,
             int n_relax = 0, double w = 0.0, bool skip_dead = false, double coarse_tol = 0.0);
// dynamic LDS (bytes) a workgroup of the kernel can use for level arrays
int mg_tail_lds_limit();
// allow launches with this much dynamic LDS (once, before the first launch)
void mg_tail_reserve_lds(int lds_bytes);

}  // namespace artss
