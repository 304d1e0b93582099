/*
 * artss_hip.h - C ABI of the MI355X (gfx950) implementation of the ARTSS
 * solver hot path.
 *
 * The reference (FireDynamics/ARTSS) has no foreign-function interface: its
 * operators are C++ classes called through IAdvection, IDiffusion,
 * IPressure, ITurbulence and ISource from the solver classes.  Each entry
 * point below is what a drop-in for one of those calls binds; the comment
 * names the reference interface it replaces (file:line in the reference's
 * src/).
 *
 * Conventions
 *  - Fields are float64 device arrays (HBM) in the reference layout: x
 *    fastest, idx = i + Nx*j + Nx*Ny*k, one ghost layer
 *    (utility/GlobalMacrosTypes.h:70).  Sizes come from artss_domain_level_info.
 *  - field_type is the reference's FieldType (utility/Mapping.h:48-51):
 *    RHO 0, U 1, V 2, W 3, P 4, T 5, NU 6; it selects the boundary data.
 *  - Every call returns 0 on success or an ARTSS_ERR_* code, with a message
 *    in artss_last_error() (per thread).  Operator calls are complete on the
 *    device when they return (the reference's sync=true).
 *  - A domain owns one HIP stream; calls on one domain must not overlap.
 */
#ifndef ARTSS_HIP_H
#define ARTSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARTSS_OK 0
#define ARTSS_ERR_ARGUMENT 1 /* bad pointer, size, level or name */
#define ARTSS_ERR_CONFIG 2   /* XML input rejected (Settings::config_error in the reference) */
#define ARTSS_ERR_HIP 3      /* HIP runtime failure */
#define ARTSS_ERR_INTERNAL 4
/* the step completed (fields advanced, counters valid) but its residual-target
 * pressure solve stopped at max_cycles with ||rhs - lap p||_2 above the
 * target, or the residual is not finite; no reference counterpart (the
 * reference's VCycleMG::pressure stops silently, VCycleMG.cpp:150-197) */
#define ARTSS_ERR_NOT_CONVERGED 5

typedef struct artss_domain artss_domain;
typedef struct artss_mg artss_mg;
typedef struct artss_solver artss_solver;
typedef struct artss_unet artss_unet;
typedef struct artss_ensemble artss_ensemble;

typedef struct {
    int N[3];               /* cells per axis including ghosts (DomainData::get_number_of_cells) */
    int n[3];               /* inner cells per axis (DomainData::get_number_of_inner_cells) */
    long long size;         /* N[0]*N[1]*N[2] (DomainData::get_size) */
    double h[3];            /* spacing (DomainData::get_spacing) */
    long long inner_cells;  /* fluid inner cells (DomainController::get_size_domain_inner_list_level_joined) */
    int n_obstacles;
} artss_level_info;

const char *artss_last_error(void);
int artss_version(void);

/* ---- domain: DomainData::init + DomainController::init (main.cpp:62-78) ---- */
int artss_domain_create(const char *xml_text, int device, artss_domain **out);
int artss_domain_destroy(artss_domain *d);
int artss_domain_levels(const artss_domain *d, int *levels);
int artss_domain_level_info(const artss_domain *d, int level, artss_level_info *info);
/* Reference index lists, for parity checks.  which: 0 inner list
 * (Domain::get_inner_list), 1 domain list, 2 obstacle cells, 3..8 domain
 * boundary patch LEFT..BACK, 10..15 face list LEFT..BACK of `obstacle`
 * (after Obstacle::remove_circular_constraints at level 0).  With out ==
 * NULL only *count is set. */
int artss_domain_list(const artss_domain *d, int which, int level, int obstacle, uint64_t *out, size_t capacity,
                      size_t *count);
int artss_domain_synchronize(artss_domain *d);

/* ---- fields: Field ctor/dtor, update_dev/update_host (field/Field.h:45-52) ---- */
int artss_field_alloc(artss_domain *d, int level, double **dev);
int artss_field_free(artss_domain *d, double *dev);
int artss_field_upload(artss_domain *d, double *dev, const double *host, size_t n);
int artss_field_download(artss_domain *d, double *host, const double *dev, size_t n);

/* ---- operators ---- */
/* DomainController::apply_boundary, domain/DomainController.cpp:145 (Multigrid.cpp:555-572) */
int artss_apply_boundary(artss_domain *d, double *field, int field_type, int level);
/* IAdvection::advect -> SLAdvect::advect, advection/SLAdvect.cpp:34 */
int artss_sl_advect(artss_domain *d, double *out, const double *in, const double *u, const double *v,
                    const double *w, int out_type);
/* IDiffusion::diffuse (both overloads, ev == NULL for the laminar one),
 * method "Jacobi" (diffusion/JacobiDiffuse.cpp:31,123),
 * "ColoredGaussSeidel" (diffusion/ColoredGaussSeidelDiffuse.cpp:38,111) or
 * "Explicit" (diffusion/ExplicitDiffuse.cpp:27,40; one step, max_iter, tol_res
 * and w unused).  `in` is modified as in the reference. */
int artss_diffuse(artss_domain *d, const char *method, double *out, double *in, const double *b, double D,
                  const double *ev, int field_type, int max_iter, double tol_res, double w, int *iterations,
                  double *residual);
/* JacobiDiffuse::JacobiStep multigrid form, diffusion/JacobiDiffuse.cpp:274 */
int artss_jacobi_step(artss_domain *d, int level, double *out, const double *in, const double *b, double alpha_x,
                      double alpha_y, double alpha_z, double beta, double dsign, double w);
/* ColoredGaussSeidelDiffuse::colored_gauss_seidel_step, diffusion/ColoredGaussSeidelDiffuse.cpp:190 */
int artss_cgs_step(artss_domain *d, int level, double *out, const double *b, double alpha_x, double alpha_y,
                   double alpha_z, double beta, double dsign, double w);
/* IPressure::divergence, interfaces/IPressure.cpp:20 */
int artss_divergence(artss_domain *d, double *out, const double *in_x, const double *in_y, const double *in_z);
/* IPressure::projection, interfaces/IPressure.cpp:125 */
int artss_projection(artss_domain *d, double *out_u, double *out_v, double *out_w, const double *in_u,
                     const double *in_v, const double *in_w, const double *in_p);
/* ITurbulence::calc_turbulent_viscosity -> ConstSmagorinsky, turbulence/ConstSmagorinsky.cpp:30 */
int artss_smagorinsky(artss_domain *d, double *ev, const double *u, const double *v, const double *w, double cs,
                      int ev_type);
/* ISource::add_source (scalar form) -> ExplicitEulerSource, source/ExplicitEulerSource.cpp:84 */
int artss_add_source(artss_domain *d, double *out, const double *s, int field_type);
/* ISource::buoyancy_force, interfaces/ISource.cpp:26 (beta, g from the XML) */
int artss_buoyancy_force(artss_domain *d, double *out, const double *T, const double *T_ambient);
/* ISourceFunction::update_source -> GaussFunction, source/GaussFunction.cpp:27 (XML temperature source) */
int artss_gauss_source(artss_domain *d, double *out, double t);

/* ---- IPressure::pressure -> VCycleMG, pressure/VCycleMG.cpp:17,142 (XML pressure section) ---- */
int artss_mg_create(artss_domain *d, artss_mg **out);
int artss_mg_destroy(artss_mg *mg);
int artss_mg_pressure(artss_mg *mg, double *p, const double *rhs, double t, int *cycles);
/* one VCycleMG::VCycleMultigrid (:215) and the residual norm of :186-195 */
int artss_mg_vcycle(artss_mg *mg, double *p, const double *rhs);
int artss_mg_residual(artss_mg *mg, const double *p, const double *rhs, double *residual);
/* artss_mg_residual without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The partial sums go to a fixed file-scope device array (1 MiB per GPU, held
 * until the process ends) instead of the domain's buffers: levels whose
 * residual kernel has more than 131,072 workgroups (64x4 threads each, e.g.
 * beyond 512x256x256 inner cells) are rejected with ARTSS_ERR_ARGUMENT.
 * All calls on one GPU share that array, so calls that overlap on one GPU
 * race.  Blocking; the same value as artss_mg_residual, bit for bit. */
int artss_mg_residual_static(artss_mg *mg, const double *p, const double *rhs, double *residual);
/* Launch schedule of the V-cycle (no reference counterpart: the reference has
 * one).  mode 0: one launch per reference operator; mode 1: launches that
 * cannot change the result folded or dropped, and the levels whose arrays
 * hold at most tail_cells cells (default 4096; 0 = none) run in one
 * single-workgroup kernel; mode 2 (default): mode 1, with each V-cycle
 * captured once into a HIP graph and replayed.  Results are bit-identical in
 * every mode. */
int artss_mg_set_schedule(artss_mg *mg, int mode, long long tail_cells);
/* first level run by the single-workgroup kernel (n_level: none) and the
 * LDS it uses for the level arrays it keeps on chip (bytes, may be NULL) */
int artss_mg_tail_info(artss_mg *mg, int *first, int *lds_bytes);

/* ---- solver: SolverController (solver/SolverController.cpp:26) + TimeIntegration ---- */
int artss_solver_create(const char *xml_text, int device, artss_solver **out);
int artss_solver_destroy(artss_solver *s);
int artss_solver_domain(artss_solver *s, artss_domain **d); /* borrowed, owned by the solver */
/* one TimeIntegration iteration: do_step(t), update_sources(t), update_data (TimeIntegration.cpp:93,141-142);
 * ARTSS_ERR_NOT_CONVERGED when a residual-target pressure solve missed its target (the step is done) */
int artss_solver_step(artss_solver *s, double t);
/* TimeIntegration::run loop for steps first..first+count-1 (t = step*dt); seconds = wall time;
 * ARTSS_ERR_NOT_CONVERGED after all steps when any of them missed its pressure target */
int artss_solver_run(artss_solver *s, long first_step, long count, double *seconds);
/* device pointer of a named field: u v w p rhs T nu_t kappa_t f_x f_y f_z S_T T_ambient u0 ... */
int artss_solver_field(artss_solver *s, const char *name, double **dev, size_t *n);
/* Analysis::save_variables_in_file (analysis/Analysis.cpp:360-408): <name>.dat (one value per cell,
 * default ostream format) and <name>_obstacle.dat ("index;value") for every field type with domain
 * boundary data, into `directory`; *files = number of files written (may be NULL) */
int artss_solver_write_dat(artss_solver *s, const char *directory, int *files);
/* last step: diffusion iterations (u, v, w, T), V-cycles, final pressure residual (-1 if not measured) */
int artss_solver_counters(artss_solver *s, int *diffusion4, int *vcycles, double *residual);
/* 0: reference pressure schedule (VCycleMG::pressure); 1: V-cycles until the residual <= tolerance */
int artss_solver_set_pressure_mode(artss_solver *s, int to_tolerance, double tolerance, int max_cycles);
/* With the residual target of artss_solver_set_pressure_mode (no reference
 * counterpart): method 0 iterates V-cycles on their own (the reference's
 * iteration); method 1 uses one V-cycle per iteration as the preconditioner
 * of a GCR iteration with one kept direction, which meets the same target
 * ||rhs - lap p||_2 in far fewer V-cycles when the V-cycle alone stalls
 * (obstacles); method 2 is method 1 with the V-cycle's levels above its
 * single-workgroup coarse levels in FP32 (smoothing, residuum, restriction,
 * prolongation), the coarse levels, the GCR iteration, residuals and norms
 * in FP64; with methods 1 and 2, coarse_max_iter > 0 sets the sweeps of the
 * preconditioner's coarse Jacobi solve (0: five; see
 * artss_solver_set_coarse_tolerance).  A GCR solve that stalls is finished
 * by method 0's V-cycles (artss_solver_fallbacks).  The final residual is
 * the true FP64 one in every method. */
int artss_solver_set_pressure_method(artss_solver *s, int method, int coarse_max_iter);
/* artss_mg_set_schedule for the solver's own pressure solver */
int artss_solver_set_mg_schedule(artss_solver *s, int fused, long long tail_cells);

/* Method 2 only (no reference counterpart): weighted-Jacobi sweeps before
 * and after the coarse correction on the FP32 levels of the preconditioner
 * (0: the XML's n_relax with the reference's effective count) and their
 * weight (0: the XML's).  The iteration still meets the same FP64 residual
 * target. */
int artss_solver_set_mixed_sweeps(artss_solver *s, int pre, int post, double weight);

/* Methods 1 and 2 (no reference counterpart): direct = 1 (the default)
 * replaces the coarse levels of GCR's preconditioner (those of the
 * single-workgroup kernel) by the exact solve on their finest level f, a
 * dense inverse assembled once when that level has at most 1536 fluid cells:
 * of its own operator where it has no obstacle cells, else of the Galerkin
 * operator R lap P (restriction, level f-1's Laplacian with its linear
 * fill, prolongation), because an obstacle grows when it is coarsened and
 * the rediscretised operator then misleads the iteration.  With direct = 0,
 * or above 1536 cells, the single-workgroup V-cycle and Jacobi coarse solve
 * run.  The iteration still meets the same FP64 residual target. */
int artss_solver_set_coarse_solve(artss_solver *s, int direct);

/* Method 2 only (no reference counterpart): n_relax of the preconditioner's
 * coarse levels (the single-workgroup kernel), whose sweeps then use the
 * weight of artss_solver_set_mixed_sweeps; 0 (the default) keeps the XML's
 * n_relax and weight. */
int artss_solver_set_mixed_tail(artss_solver *s, int n_relax);
/* Stopping test of the coarse solve inside the GCR preconditioner of
 * methods 1 and 2 (no reference counterpart; the reference's coarse solve
 * stops at the XML's absolute tol_res).  coarse_tol 0: that absolute
 * tolerance; 0 < coarse_tol < 1: stop once the coarse residual has fallen by
 * that factor from its start, whatever the scale of the right-hand side;
 * < 0 (the default): no test, a fixed number of sweeps (the coarse_max_iter
 * of artss_solver_set_pressure_method, 0: five).  An absolute tolerance makes
 * the preconditioner weaker as the GCR residual shrinks, which stalls GCR on
 * the 256x128x128 plume from step 709 on; a fixed count keeps the
 * preconditioner one linear operator. */
int artss_solver_set_coarse_tolerance(artss_solver *s, double coarse_tol);
/* Steps since creation whose GCR pressure solve (methods 1, 2) stopped
 * falling for GCR's stall window and was finished by the reference's V-cycle
 * iteration within the same cycle budget (no reference counterpart). */
int artss_solver_fallbacks(artss_solver *s, long long *steps);

/* No reference counterpart: the cells per thread along z (1, 2 or 4) that the
 * autotuner picked for the fused level kernels of levels with this many
 * cells (array size with ghosts), or 0 when no level of that size has been
 * tuned yet (the mixed pressure path tunes its FP32 levels on first use). */
int artss_tuned_cells_per_thread(long long level_size);

/* Start of the residual-target GCR pressure iteration (methods 1 and 2; no
 * reference counterpart: the reference's iteration starts from the previous
 * step's p, mode 0).  With p_n the previous solution, p_prev the one before
 * and d = p_n - p_prev: mode 1 starts from p_n + d, mode 2 from p_n + d + s *
 * net(c (rhs - lap(p_n + d)) / s, d / s) with s = rms(d) over the fluid cells and
 * c = 1 / (2/dx^2 + 2/dy^2 + 2/dz^2), `net` created on the solver's domain.
 * The iteration runs to the same residual target from any start.  `record`
 * keeps each step's network input and its target (p_{n+1} - p_n - d) / s.
 * Mode 2 and `record` need inner box sides that are multiples of 8. */
int artss_solver_set_warm_start(artss_solver *s, int mode, artss_unet *net, int record);
/* The same for the starts that need no network, with `depth` for mode 3:
 * p_n + sum_s alpha_s d_s over the last `depth` (1 .. 8) increments d_s of
 * the solution, the combination closest to the solution in the energy norm
 * (Galerkin projection onto the previous steps' increments; its error in
 * that norm is never larger than that of mode 1, alpha = (1, 0, ...)).  No
 * reference counterpart.  A restore or a load clears the kept increments. */
int artss_solver_set_warm_start_depth(artss_solver *s, int mode, int depth);

/* Step graph (no reference counterpart; for ensembles of small domains,
 * where a step's launches and host round trips outweigh its kernels).
 * After one eager step that records how many iterations each residual-target
 * loop took, artss_solver_step captures the whole step in one HIP graph
 * (budget: that count + `margin` iterations per loop, behind the loop's
 * device test) and replays it, reading one outcome per loop.  The graph first
 * copies the fields aside; a replay in which a loop would have needed more
 * than its budget is rolled back and the step runs eagerly and is captured
 * again.  The fields are the eager step's bit for bit.  Needs the
 * residual-target GCR pressure solve, warm start 0 or 1 without recording,
 * and colored Gauss-Seidel diffusion with an idempotent fill; otherwise the
 * steps run eagerly and `note` (valid until the next call) says why. */
int artss_solver_set_step_graph(artss_solver *s, int on, int margin);  /* margin < 0: budgets below the
                                                                         recorded counts (rollback tests) */
int artss_solver_step_graph_stats(artss_solver *s, long long *replays, long long *rollbacks, long long *captures,
                                  int *active, const char **note);
/* copies the last recorded sample to device buffers x [V][2] and t [V] (V
 * inner box cells, x fastest); *available = 0 when none is recorded yet */
int artss_solver_warm_sample(artss_solver *s, float *x, float *t, int *available);
/* ||rhs - lap p||_2 at the start of the last step's GCR pressure iteration */
int artss_solver_initial_residual(artss_solver *s, double *residual);

/* ---- checkpoint and rollback (dataAssimilation/FieldIO.cpp:44-100 writes
 * u, v, w, p, T, C per time to HDF5, DataAssimilation.cpp:56-161 rolls a run
 * back to such a state).  The state between two steps is every field of the
 * solver plus the warm start's previous pressure; restoring it continues the
 * run bit for bit.  Snapshots stay in HBM (slot ids, freed by
 * artss_solver_drop_snapshot or with the solver); save/load write one file
 * of this library's own format (magic, version, cell counts, time, dt, field
 * names, raw FP64 arrays; HDF5 is not supported) and load checks
 * that the grid matches. */
int artss_solver_snapshot(artss_solver *s, double t, int *slot);
int artss_solver_restore(artss_solver *s, int slot, double *t);
int artss_solver_drop_snapshot(artss_solver *s, int slot);
int artss_solver_save(artss_solver *s, const char *path, double t);
int artss_solver_load(artss_solver *s, const char *path, double *t);

/* ---- ensembles (no reference counterpart: the reference advances one
 * domain per process).  Independent solvers of any sizes, each on its own
 * stream, advanced together: artss_ensemble_step runs Solver::step(step *
 * dt_m) for every member m on `threads` host threads, so the members'
 * kernels overlap on the GPU.  Members are not owned; they must outlive the
 * ensemble and not be stepped elsewhere during artss_ensemble_step.  An
 * error in any member is reported after all members are done. */
int artss_ensemble_create(int threads, artss_ensemble **out);
int artss_ensemble_destroy(artss_ensemble *e);
int artss_ensemble_add(artss_ensemble *e, artss_solver *member);
int artss_ensemble_step(artss_ensemble *e, long step);

/* ---- raw device buffers on a domain's device (any element type) ---- */
int artss_buffer_alloc(artss_domain *d, size_t bytes, void **dev);
int artss_buffer_free(artss_domain *d, void *dev);
int artss_buffer_upload(artss_domain *d, void *dev, const void *host, size_t bytes);
int artss_buffer_download(artss_domain *d, void *host, const void *dev, size_t bytes);

/* ---- warm-start U-Net for the pressure solve (no reference counterpart) ----
 * A 3-D convolutional U-Net (levels with 16, 32, 64, 64 channels, 3x3x3
 * convolutions, BF16 weights and activations, FP32 accumulation on the
 * matrix cores) on the domain's stream.  Tensors are channels-last FP32
 * device arrays over an X x Y x Z voxel block (sides multiples of 8):
 * input [V][2], output and target [V].  Parameters are FP32 in layer order
 * in, e0, e1, e2, e3, d2, d1, d0, out, each W[27][c_in][c_out] then b[c_out]. */
int artss_unet_create(artss_domain *d, unsigned long long seed, artss_unet **out);
int artss_unet_destroy(artss_unet *u);
int artss_unet_param_count(size_t *n);
int artss_unet_set_params(artss_unet *u, const float *host);
int artss_unet_get_params(artss_unet *u, float *host);
int artss_unet_forward(artss_unet *u, int X, int Y, int Z, const float *x, float *y);
/* mean squared error against target and its gradient (host `grads` may be NULL) */
int artss_unet_gradients(artss_unet *u, int X, int Y, int Z, const float *x, const float *target, float *grads,
                         float *loss);
/* one Adam step (beta 0.9, 0.999) on the gradients of the last artss_unet_gradients call */
int artss_unet_adam_step(artss_unet *u, float lr);

/* ---- live kernel timing (no reference counterpart; the reference only
 * prints "Global Time", TimeIntegration.cpp:165-183) ----
 * Arms HIP-event timing of every launch of one kernel at one level, tag
 * "<kernel>@<level>" (kernels: jacobi_mg cgs_mg jacobi_diffuse cgs_diffuse
 * sl_advect residuum residual restrict prolongate bc_faces bc_domain);
 * read returns the launches seen since arming and their summed duration. */
int artss_profile_arm(artss_domain *d, const char *tag);
int artss_profile_read(artss_domain *d, long long *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
