// Host-side launchers for the HIP kernels of the solver hot path.
// Every launcher enqueues on the given stream and returns immediately; the
// caller owns synchronisation.  All arrays are device pointers in the
// reference layout (device.h).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace artss {
namespace k {

// ---- element-wise over the whole array (Field.h:64-135) ----
// gate (optional): nothing runs while *gate != 0 (IterGate::done)
void copy(hipStream_t s, double *dst, const double *src, long long n, const int *gate = nullptr);
// x[cells[r]] = sum_j M[r * n + j] b[cells[j]] for r < n (n <= kMaxDenseRows:
// the gathered b is staged in 16 KiB of LDS per workgroup)
constexpr int kMaxDenseRows = 2048;
void dense_apply(hipStream_t s, int n, const double *M, const uint32_t *cells, const double *b, double *x);
// dst0[q] = dst1[q] = src[q] for q < fields (1..4), n elements each, one pass
struct CopyFan {
    const double *src[4];
    double *dst0[4];
    double *dst1[4];
    int fields;
};
void copy_fan(hipStream_t s, const CopyFan &f, long long n);
void fill(hipStream_t s, double *dst, double v, long long n);
void fill32(hipStream_t s, float *dst, float v, long long n);
void scale(hipStream_t s, double *dst, double v, long long n);           // dst *= v
void add(hipStream_t s, double *dst, const double *src, long long n);    // dst += src
void scale_to(hipStream_t s, double *dst, const double *src, double v, long long n);  // dst = src * v
void scale_to_dev(hipStream_t s, double *dst, const double *src, const double *v, long long n);  // dst = src * *v
// dst[q] := src[q] for q < count arrays of n values; the pointer arrays are in device memory
void copy_many(hipStream_t s, double *const *dst, const double *const *src, int count, long long n);
// copy on the domain list (inner + ghost shell, obstacle cells untouched), JacobiDiffuse.cpp:85-99
void copy_domain(hipStream_t s, const LevelDesc &L, double *dst, const double *src);

// ---- boundary fill (Multigrid.cpp:555-572) ----
void bc_faces(hipStream_t s, const FaceEntry *e, long long n, const FaceParam *p, double *f, const int *gate = nullptr);
void bc_domain(hipStream_t s, const LevelDesc &L, const DomainBC &bc, double *f, const int *gate = nullptr);
// faces (one wave) and ghosts in one launch; only for a fill none of whose
// writes reads a cell another write changes
void bc_merged(hipStream_t s, const LevelDesc &L, const DomainBC &bc, const FaceEntry *e, long long n_faces,
               const FaceParam *p, double *f);

// ---- SLAdvect::advect (SLAdvect.cpp:71-163) ----
void sl_advect(hipStream_t s, const LevelDesc &L, double dt, double *out, const double *in, const double *u,
               const double *v, const double *w);

// ---- Jacobi diffusion (JacobiDiffuse.cpp) with fused update-size partials ----
// plain level-0 form divides by rb (:241-249), turbulent form (:345-362) when ev != nullptr.
// partial[b] = sum over the block of (rb*(out-in))^2 (:72-77, :162-172)
void jacobi_diffuse_step(hipStream_t s, const LevelDesc &L, double *out, const double *in, const double *b,
                         double ax, double ay, double az, double rb, double D, const double *ev, double dt,
                         double w, double *partial);
// ISource.cpp:63-114: out += dt * (viscous dissipation) on inner cells
void dissipate(hipStream_t s, const LevelDesc &L, double *out, const double *u, const double *v, const double *w,
               double nu, double dt);
// ExplicitDiffuse.cpp:49-121: one explicit Euler step on inner cells (ev == nullptr: plain form)
void explicit_diffuse(hipStream_t s, const LevelDesc &L, double *out, const double *in, double D, const double *ev,
                      double dt);
// multigrid form, multiplies by beta (:292-300)
void jacobi_step_mg(hipStream_t s, const LevelDesc &L, double *out, const double *in, const double *b, double ax,
                    double ay, double az, double beta, double dsign, double w);
// first sweep with the initial copy folded in: dst = update(src) on inner cells, src elsewhere
void jacobi_mg_first(hipStream_t s, const LevelDesc &L, double *dst, const double *src, const double *b, double ax,
                     double ay, double az, double beta, double dsign, double w);

// ---- colored Gauss-Seidel (ColoredGaussSeidelDiffuse.cpp:190-341), one color per launch ----
// ev == nullptr: constant ax,ay,az,beta; else per-cell coefficients from D, ev, dt
void cgs_color(hipStream_t s, const LevelDesc &L, int color, double *out, const double *b, double ax, double ay,
               double az, double beta, double dsign, double w, double D, const double *ev, double dt,
               const int *gate = nullptr);
// (rb*(out-in))^2 partials over inner cells, ColoredGaussSeidelDiffuse.cpp:77-82, :147-157
void update_norm_partials(hipStream_t s, const LevelDesc &L, const double *out, const double *in, double rb,
                          double D, const double *ev, double dt, double *partial, const int *gate = nullptr);
// coefficients of one colored Gauss-Seidel iteration: constant ax, ay, az,
// beta = 1/rb (ev == nullptr), else per cell from D, ev, dt and 1/h
struct CgsCoef {
    double ax, ay, az, beta, rb, dsign, w, D, dt, rdx, rdy, rdz;
};
// one whole iteration (even sweep, odd sweep, fill of the result, norm
// partials) out of place from src into dst in one launch; the fill must be
// idempotent (Domain::bc_idempotent).  cgs_iter_partials: the partial sums
// it writes.
int cgs_iter_partials(const LevelDesc &L);
void cgs_iter(hipStream_t s, const LevelDesc &L, const BCView &V, const CgsCoef &c, const double *ev, const double *b,
              const double *src, double *dst, double *partial, const int *gate);
// exchanges the contents of two arrays of n values
void swap_arrays(hipStream_t s, double *a, double *b, long long n);

// ---- multigrid (VCycleMG.cpp) ----
void residuum(hipStream_t s, const LevelDesc &L, double *out, const double *x, const double *b);   // :346-354
void residual_partials(hipStream_t s, const LevelDesc &L, const double *x, const double *b,
                       double *partial);                                                          // :186-193
void restrict_(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, double *coarse, const double *fine);
void prolongate(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, double *fine, const double *coarse);
// coarse = restrict(fine) on coarse inner cells and coarse_err = 0 on the whole coarse array
// (instantiated for double/double/double, float/float/float and float in -> double out)
template <typename Tin, typename Tout, typename Terr>
void restrict_fill(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, Tout *coarse, const Tin *fine,
                   Terr *coarse_err);
// fine += prolongate(coarse) on the cells the prolongation writes
void prolongate_add(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, double *fine, const double *coarse);

// ---- fused operator + boundary fill, for idempotent fills (device.h BCView) ----
// Iterates of type T (double: the reference's arithmetic; float: the mixed
// V-cycle), right-hand sides / coarse errors of type TB / Tc.
void copy_bc(hipStream_t s, const LevelDesc &L, const BCView &V, double *dst, const double *src);
// restrict_fill<float, float, float> into rhs followed by the coarse level's
// first sweep from zero into err with its fill (jacobi_bc with src == nullptr),
// in one launch; the values are those of the two launches
void restrict_sweep0(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, const BCView &V, float *rhs, float *err,
                     const float *fine, double ax, double ay, double az, double beta, double dsign, double w);
// autotuner of the fused level kernels' cells per thread along z (1, 2 or 4)
// for a level size, timed on scratch arrays; call outside stream capture
void tune_level(hipStream_t s, const LevelDesc &L, const BCView &V);
int tuned_kz(long long level_size);  // 0: not tuned (the size rule applies)
// 0: autotuned choice; 1, 2 or 4: every fused level kernel uses that many
// cells per thread along z (results are the same for every choice)
void set_forced_kz(int kz);
// src == nullptr (first sweeps only): the sweep starts from zero everywhere
template <typename T, typename TB>
void jacobi_bc(hipStream_t s, const LevelDesc &L, const BCView &V, bool first, T *dst, const T *src, const TB *b,
               double ax, double ay, double az, double beta, double dsign, double w);
template <typename T, typename TB>
void residuum_faces(hipStream_t s, const LevelDesc &L, const BCView &V, T *out, const T *x, const TB *b);
template <typename Tf, typename Tc>
void prolongate_add_bc(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, const BCView &V, Tf *dst,
                       const Tf *fine, const Tc *coarse);
template <typename T>
void fill_bc(hipStream_t s, const LevelDesc &L, const BCView &V, T *f);

// ---- reductions ----
int partial_count(const LevelDesc &L);  // number of partial slots a level-wide kernel writes
void reduce_sum(hipStream_t s, const double *partial, int n, double *out);  // fixed-order sum
// Loop state of an iteration whose test runs on the device, so that several
// iterations can be enqueued without a readback each: kernels given &done do
// nothing once it is set, and the state after the batch is that of a loop
// that tested after every iteration.
struct IterGate {
    double sum;    // the last sum reduced (the residual is its square root)
    double tol;
    int count;     // iterations that ran since gate_open
    int max_iter;
    int done;      // 1: the loop would have stopped
};
void gate_open(hipStream_t s, IterGate *g, double tol, int max_iter);
// log := the gate's count, done and sqrt(sum)
void log_gate(hipStream_t s, const IterGate *g, LoopLog *log);
// the end of an out-of-place loop of g->count iterations alternating between
// out (odd iterations write it) and in: after none, out := in; after an even
// count the arrays' contents are exchanged, so out holds the last iterate
void alternate_finish(hipStream_t s, const IterGate *g, double *out, double *in, long long n);
// reduce_sum into g->sum, then count the iteration and set done as the
// loop `while (sqrt(sum) > tol && count < max_iter)` decides; nothing once done
void reduce_gate(hipStream_t s, const double *partial, int n, IterGate *g);

// ---- DynamicSmagorinsky (turbulence/DynamicSmagorinsky.cpp) ----
constexpr int kDynMax = 12;
struct DynArrays {
    double *p[kDynMax];
};
struct DynConstArrays {
    const double *p[kDynMax];
};
void dyn_filter(hipStream_t s, const LevelDesc &L, const DynArrays &out, const DynConstArrays &in, int count);
void dyn_products(hipStream_t s, const LevelDesc &L, const double *uf, const double *vf, const double *wf,
                  const DynArrays &prod);
void dyn_strain(hipStream_t s, const LevelDesc &L, const double *uf, const double *vf, const double *wf,
                const DynArrays &out, double *sbar);
void dyn_coef(hipStream_t s, const LevelDesc &L, double c2, double a2, const DynConstArrays &filt,
              const DynConstArrays &prod, const DynConstArrays &sf, double *cs);
// the in-place average in index order as (n_x + 2 n_y + 4 n_z) wavefront launches
void dyn_average(hipStream_t s, const LevelDesc &L, double delta_s, double nu, double *cs, const double *sbar,
                 double *ev);

// ---- IPressure (IPressure.cpp) ----
void divergence(hipStream_t s, const LevelDesc &L, double *out, const double *x, const double *y, const double *z);
void projection(hipStream_t s, const LevelDesc &L, double *ou, double *ov, double *ow, const double *iu,
                const double *iv, const double *iw, const double *p);

// ---- ConstSmagorinsky (ConstSmagorinsky.cpp:60-80) ----
void smagorinsky(hipStream_t s, const LevelDesc &L, double *ev, const double *u, const double *v, const double *w,
                 double cs2_delta2);

// ---- sources ----
void add_source(hipStream_t s, const LevelDesc &L, double *out, const double *src, double dt);   // inner list
void buoyancy(hipStream_t s, const LevelDesc &L, double *out, const double *T, const double *Ta, double beta,
              double g);                                                                           // domain list

}  // namespace k
}  // namespace artss
