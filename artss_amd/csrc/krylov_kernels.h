// Launchers of the Krylov pressure-solve kernels (krylov_kernels.hip).  Partial-sum
// buffers hold q * blocks + block for q = 0.. of each kernel's sums.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace artss {
namespace kk {

int blocks_inner(const LevelDesc &L);  // partial slots per sum of the inner-grid kernels
int blocks_full(const LevelDesc &L);   // ... of the whole-array kernels

// r = b - lap(x) on inner cells (x must carry its boundary fill); 1 sum: (r, r)
void residual_vec(hipStream_t s, const LevelDesc &L, double *r, const double *x, const double *b, double *partial);
// w = lap(z) on inner cells; 4 sums: (w,w), (w,wp), (r,w), (r,wp)
template <typename Tz>
void lap_dots(hipStream_t s, const LevelDesc &L, double *w, const Tz *z, const double *wp, const double *r,
              double *partial);
// GCR coefficients from the 4 sums: coef[0] beta, coef[1] 1/|w - beta wp|, coef[2] alpha
void gcr_coef(hipStream_t s, const double *partial, int blocks, double *coef);
// the GCR step over the whole array; 1 sum: (r, r)
template <typename Tz>
void gcr_update(hipStream_t s, const LevelDesc &L, const double *coef, double *x, double *r, const Tz *z, double *zp,
                const double *w, double *wp, double *partial);
// out[0] = fixed-order sum of n partials
void sum(hipStream_t s, const double *partial, int n, double *out);

}  // namespace kk
}  // namespace artss
