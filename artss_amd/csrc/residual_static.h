// Residual norm over a fixed-size, file-scope device scratch array: no
// device allocation per call (residual_static.hip).
#pragma once

namespace artss {

class Domain;

// partial sums the static scratch holds: one per workgroup of a level-wide
// kernel (64x4 threads), e.g. 16,384 at 256x128x128 inner cells, 131,072 at
// 512x256x256
constexpr int kStaticMaxPartials = 1 << 17;

// looks up the scratch array's device address on `device` (once per device)
void static_scratch_setup(int device);
// ||rhs - lap p||_2 over the fluid inner cells of level 0, as
// VCycleMG::residual computes it; blocking
double residual_static(Domain &d, const double *p, const double *rhs);

}  // namespace artss
