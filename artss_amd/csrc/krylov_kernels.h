// Launchers of the Krylov pressure-solve kernels (krylov_kernels.hip).  Partial-sum
// buffers hold q * blocks + block for q = 0.. of each kernel's sums.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace artss {
namespace kk {

// Convergence gate of a GCR solve in device memory: iterations enqueued
// after the one that met the tolerance change nothing, so several can be
// enqueued without waiting for each residual.
struct GcrGate {
    double rr;    // (r, r) after the last iteration that ran
    double tol2;  // tolerance squared
    int count;    // iterations that ran since gate_open
    int done;     // 1 once rr <= tol2
    double rr0;   // (r, r) at gate_start
};

int blocks_inner(const LevelDesc &L);  // partial slots per sum of the inner-grid kernels
int blocks_full(const LevelDesc &L);   // ... of the whole-array kernels
int blocks_dots(const LevelDesc &L);   // ... of lap_dots

// r = b - lap(x) on inner cells (x must carry its boundary fill); 1 sum: (r, r)
void residual_vec(hipStream_t s, const LevelDesc &L, double *r, const double *x, const double *b, double *partial);
// w = lap(z), wp = lap(zp) on inner cells (not stored); 4 sums: (w,w), (w,wp), (r,w), (r,wp)
template <typename Tz>
void lap_dots(hipStream_t s, const LevelDesc &L, const Tz *z, const Tz *zp, const double *r, double *partial);
// GCR coefficients from the 4 sums: coef[0] beta, coef[1] 1/|w - beta wp|, coef[2] alpha
// (gate: nothing when closed; nullptr: always)
void gcr_coef(hipStream_t s, const double *partial, int blocks, double *coef, const GcrGate *gate);
// the GCR step over the whole array (w = lap(z) and wp = lap(zp)
// recomputed); 1 sum: (r, r).  The new kept direction goes to zp_out, a
// buffer other than zp.
template <typename Tz>
void gcr_update(hipStream_t s, const LevelDesc &L, const double *coef, double *x, double *r, const Tz *z,
                const Tz *zp, Tz *zp_out, double *partial, const GcrGate *gate);
// out[0] = fixed-order sum of n partials
void sum(hipStream_t s, const double *partial, int n, double *out);
// the same sum into gate->rr, counting the iteration and closing the gate
// when rr <= tol2; nothing when the gate is closed
void sum_gate(hipStream_t s, const double *partial, int n, GcrGate *gate);
// rr = 0, count = 0, done = 0, tol2 = tol2
void gate_open(hipStream_t s, GcrGate *gate, double tol2);
// rr = the fixed-order sum of n partials (as sum), count = 0, done = rr <= tol2
void gate_start(hipStream_t s, const double *partial, int n, GcrGate *gate, double tol2);
// log: the gate's count, t = sqrt of the sum of n partials (in the order of
// k::reduce_sum), ok = the gate closed and t <= tol
void gate_confirm(hipStream_t s, const double *partial, int n, const GcrGate *gate, double tol, LoopLog *log);

}  // namespace kk
}  // namespace artss

namespace artss {
namespace kk {

// ---- warm start of the pressure solve (solver.cpp), p_n the previous
// solution (carrying its fill), p_prev the one before, d = p_n - p_prev,
// s = rms(d) over the fluid cells of level 0, inner box volumes x fastest
// *sum = sum of d^2 over the fluid cells
void warm_sum(hipStream_t s, const LevelDesc &L, const double *p, const double *p_prev, double *partial,
              double *sum);
// network input [v][2]: c (b - lap(p + d)) / s and d / s, zero off the fluid
void warm_pack(hipStream_t s, const LevelDesc &L, const double *p, const double *b, const double *p_prev,
               const double *sum, double fluid_cells, double c, float *x);
// p += d (+ s y[v] when y is given) on the fluid cells; nothing when d == 0
void warm_apply(hipStream_t s, const LevelDesc &L, const double *sum, double fluid_cells, const float *y,
                const double *p_prev, double *p);
// training target [v]: (p_new - p_n - d) / s, zero off the fluid
void warm_target(hipStream_t s, const LevelDesc &L, const double *sum, double fluid_cells, const double *p_new,
                 const double *p_start, const double *p_prev, float *t);

// ---- projection start (solver.cpp): p_n + sum_s alpha_s D_s over the last
// increments D_s of the solution kept in a ring of slots, alpha from the
// Galerkin condition (D_s, b - lap(p_n) - sum_t alpha_t W_t) = 0 with W_t the
// image of D_t: the combination closest to the solution in the energy norm
constexpr int kProjMax = 8;  // slots
struct ProjRing {
    double *d[kProjMax];  // increments, fluid cells
    double *w[kProjMax];  // their images, fluid cells
};
// slot `fresh` := p - p_prev and its image; partial sums (blocks_inner each)
// of (D_fresh, W_s) for q = s and (D_s, b - lap p) for q = kProjMax + s,
// over the slots whose bit is set in `valid` (which includes `fresh`)
void proj_update(hipStream_t s, const LevelDesc &L, const double *p, const double *p_prev, const double *b,
                 const ProjRing &ring, int fresh, unsigned valid, double *partial);
// G (kProjMax x kProjMax, by slot) gets row and column `fresh` of
// -(D_s, W_t); alpha (by slot) solves the projection over the valid slots
void proj_solve(hipStream_t s, const double *partial, int blocks, int fresh, unsigned valid, double *G,
                double *alpha);
// p += sum_s alpha_s D_s on the fluid cells
void proj_apply(hipStream_t s, const LevelDesc &L, const ProjRing &ring, int fresh, unsigned valid,
                const double *alpha, double *p);

}  // namespace kk
}  // namespace artss
