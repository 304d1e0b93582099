/* Oracle solver drivers.  TEST INFRASTRUCTURE ONLY.
 * Restates solver/NSSolver.cpp:50-124, solver/NSTempTurbSolver.cpp:94-248,
 * solver/NSTempSolver.cpp:69-198, the do_step of solver/NSTurbSolver.cpp,
 * DiffusionSolver.cpp, AdvectionDiffusionSolver.cpp, DiffusionTurbSolver.cpp
 * and AdvectionSolver.cpp, the concentration step of NSTempConSolver.cpp and
 * NSTempTurbConSolver.cpp (whose concentration advection the reference never
 * sets up: here it is the SL advection every other field uses),
 * solver/SolverController.cpp:26-43,229-307, field/FieldController.cpp:101-185
 * and one iteration of TimeIntegration.cpp:86-149. */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_sim {
    const or_domain *D;
    or_sim_params p;
    or_mg *mg;
    size_t n;
    double *u, *v, *w, *u0, *v0, *w0, *ut, *vt, *wt;
    double *p_, *rhs, *T, *T0, *Tt, *Ta, *nu_t, *kappa_t, *fx, *fy, *fz, *S_T, *spatial;
    double *lin[3]; /* AdvectionSolver's constant velocity fields */
    double *C, *C0, *Ct, *S_C, *gamma_t, *con_spatial; /* concentration solvers */
    int counters[5];
};

static double *alloc0(size_t n) { return (double *)calloc(n, sizeof(double)); }
static void cp(double *dst, const double *src, size_t n) { memcpy(dst, src, n * sizeof(double)); }

static int has_dir(const int *dir, int a) { return dir[a] != 0; }

/* FieldController::couple_vector / couple_scalar, FieldController.cpp:151-186 */
static void couple_vector(or_sim *s) {
    cp(s->u0, s->u, s->n); cp(s->ut, s->u, s->n);
    cp(s->v0, s->v, s->n); cp(s->vt, s->v, s->n);
    cp(s->w0, s->w, s->n); cp(s->wt, s->w, s->n);
}
static void couple_scalar(or_sim *s) {
    cp(s->T0, s->T, s->n);
    cp(s->Tt, s->T, s->n);
}
static void couple_con(or_sim *s) {
    cp(s->C0, s->C, s->n);
    cp(s->Ct, s->C, s->n);
}
static int is_temp(int solver) { return solver == 1 || solver == 3 || solver == 9 || solver == 10; }
static int is_con(int solver) { return solver == 9 || solver == 10; }

static int diffuse(or_sim *s, double *out, double *in, const double *b, double Dc, const double *ev, int ft) {
    const or_sim_params *p = &s->p;
    if (p->diff_type == 0)
        return or_jacobi_diffuse(s->D, out, in, b, Dc, ev, p->dt, p->diff_max_iter, p->diff_tol, p->diff_w, ft, NULL);
    if (p->diff_type == 2) {
        or_explicit_diffuse(s->D, out, in, Dc, ev, p->dt, ft);
        return 1;
    }
    return or_cgs_diffuse(s->D, out, in, b, Dc, ev, p->dt, p->diff_max_iter, p->diff_tol, p->diff_w, ft, NULL);
}

/* SolverController::update_sources, SolverController.cpp:263-307 */
static void update_sources(or_sim *s, double t) {
    const or_sim_params *p = &s->p;
    if (p->has_force) {
        double *f[3] = {s->fx, s->fy, s->fz};
        for (int a = 0; a < 3; a++)
            if (has_dir(p->force_dir, a)) or_buoyancy(s->D, f[a], s->T, s->Ta, p->beta, p->g);
    }
    if (is_temp(p->solver) && p->has_temp_source) { /* GaussFunction::update_source, GaussFunction.cpp:27-40 */
        double tv = tanh(t / p->tau);
        cp(s->S_T, s->spatial, s->n);
        for (size_t i = 0; i < s->n; i++) s->S_T[i] *= tv;
    }
    if (is_con(p->solver)) { /* NSTempTurbConSolver::update_source, :358-361 */
        cp(s->S_C, s->con_spatial, s->n);
        if (p->con_fct == 1) {
            double tv = tanh(t / p->con_tau);
            for (size_t i = 0; i < s->n; i++) s->S_C[i] *= tv;
        }
    }
}

/* FieldController::update_data, FieldController.cpp:119-135 */
static void update_data(or_sim *s) {
    cp(s->u0, s->u, s->n); cp(s->v0, s->v, s->n); cp(s->w0, s->w, s->n);
    cp(s->ut, s->u, s->n); cp(s->vt, s->v, s->n); cp(s->wt, s->w, s->n);
    cp(s->T0, s->T, s->n); cp(s->Tt, s->T, s->n);
    if (s->C) couple_con(s);
}

/* Cube source (source/Cube.cpp:35-80) over the cells whose centres lie in the
 * box, by the obstacles' index rule, clipped to the inner cells, obstacle
 * cells left zero (the reference computes the box's end index from the
 * domain's end coordinate, which selects no cell or runs past the array) */
static void cube_field(const or_domain *D, double *out, const double *c) {
    size_t s[3], e[3];
    for (int a = 0; a < 3; a++) {
        long lo = (long)round((c[a] - D->x0[a]) / D->h[0][a]) + 1;
        long hi = (long)round((c[3 + a] - D->x0[a]) / D->h[0][a]);
        if (lo < 1) lo = 1;
        if (hi > (long)D->n[0][a]) hi = (long)D->n[0][a];
        s[a] = (size_t)lo;
        e[a] = hi < lo ? 0 : (size_t)hi;
    }
    const ivec *inner = &D->inner[0];
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    for (size_t l = 0; l < inner->n; l++) {
        const size_t idx = inner->d[l];
        const size_t k = idx / (Nx * Ny), j = (idx / Nx) % Ny, i = idx % Nx;
        if (i >= s[0] && i <= e[0] && j >= s[1] && j <= e[1] && k >= s[2] && k <= e[2]) out[idx] = c[6];
    }
}

or_sim *or_sim_create(const or_domain *D, const or_sim_params *p, const double *u_ic, const double *v_ic,
                      const double *w_ic, const double *T_ic, const double *p_ic) {
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    s->D = D;
    s->p = *p;
    s->n = D->size[0];
    const size_t n = s->n;
    double **all[] = {&s->u, &s->v, &s->w, &s->u0, &s->v0, &s->w0, &s->ut, &s->vt, &s->wt, &s->p_, &s->rhs,
                      &s->T, &s->T0, &s->Tt, &s->Ta, &s->nu_t, &s->kappa_t, &s->fx, &s->fy, &s->fz, &s->S_T,
                      &s->spatial};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) *all[i] = alloc0(n);
    /* the drivers with a pressure solve */
    s->mg = (p->solver <= 1 || p->solver == 3 || p->solver == 4 || is_con(p->solver)) ? or_mg_create(D, &p->mg)
                                                                                         : NULL;
    if (is_con(p->solver)) {
        double **con[] = {&s->C, &s->C0, &s->Ct, &s->S_C, &s->gamma_t, &s->con_spatial};
        for (size_t i = 0; i < sizeof(con) / sizeof(con[0]); i++) *con[i] = alloc0(n);
        if (p->con_fct == 1) or_gauss_spatial(D, s->con_spatial, p->con_gauss);
        else cube_field(D, s->con_spatial, p->con_cube);
    }
    /* set_up_fields: initial conditions on the inner list, SolverController.cpp:95-205 */
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; l++) {
        size_t i = inner->d[l];
        if (u_ic) s->u[i] = u_ic[i];
        if (v_ic) s->v[i] = v_ic[i];
        if (w_ic) s->w[i] = w_ic[i];
        if (T_ic) s->T[i] = T_ic[i];
        if (p_ic) s->p_[i] = p_ic[i];
    }
    const int temp = is_temp(p->solver);
    if (temp && p->has_force) /* force_source: T_ambient.set_value, :229-236 */
        for (size_t i = 0; i < n; i++) s->Ta[i] = p->T_ambient;
    if (temp && p->has_temp_source) or_gauss_spatial(D, s->spatial, p->gauss);
    if (p->solver == 8) /* AdvectionSolver.cpp:16-34, Field(FieldType, u_lin) */
        for (int a = 0; a < 3; a++) {
            s->lin[a] = alloc0(n);
            for (size_t i = 0; i < n; i++) s->lin[a][i] = p->u_lin[a];
        }
    /* FieldController::set_up_boundary, FieldController.cpp:101-112 */
    or_apply_bc(D, s->u, OR_U, 0);
    or_apply_bc(D, s->v, OR_V, 0);
    or_apply_bc(D, s->w, OR_W, 0);
    or_apply_bc(D, s->p_, OR_P, 0);
    or_apply_bc(D, s->T, OR_T, 0);
    if (s->C) or_apply_bc(D, s->C, OR_RHO, 0);
    if (temp) or_apply_bc(D, s->Ta, OR_T, 0);
    else { /* T_ambient of a solver without temperature is never read */ }
    update_sources(s, 0);
    update_data(s);
    return s;
}

void or_sim_free(or_sim *s) {
    if (!s) return;
    double *all[] = {s->u, s->v, s->w, s->u0, s->v0, s->w0, s->ut, s->vt, s->wt, s->p_, s->rhs,
                     s->T, s->T0, s->Tt, s->Ta, s->nu_t, s->kappa_t, s->fx, s->fy, s->fz, s->S_T, s->spatial};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) free(all[i]);
    for (int a = 0; a < 3; a++) free(s->lin[a]);
    double *con[] = {s->C, s->C0, s->Ct, s->S_C, s->gamma_t, s->con_spatial};
    for (size_t i = 0; i < sizeof(con) / sizeof(con[0]); i++) free(con[i]);
    or_mg_free(s->mg);
    free(s);
}

/* ITurbulence::calc_turbulent_viscosity (ConstSmagorinsky.cpp:30-87 or
 * DynamicSmagorinsky.cpp:172-393) on nu_t */
static void turbulence(or_sim *s) {
    const or_sim_params *p = &s->p;
    if (p->turb_dynamic)
        or_dyn_smagorinsky(s->D, s->nu_t, s->u, s->v, s->w, p->nu);
    else
        or_smagorinsky(s->D, s->nu_t, s->u, s->v, s->w, p->cs, OR_NU);
}

static void do_step_ns(or_sim *s, double t) {
    const or_domain *D = s->D;
    const or_sim_params *p = &s->p;
    or_sl_advect(D, s->u, s->u0, s->u0, s->v0, s->w0, p->dt, OR_U);
    or_sl_advect(D, s->v, s->v0, s->u0, s->v0, s->w0, p->dt, OR_V);
    or_sl_advect(D, s->w, s->w0, s->u0, s->v0, s->w0, p->dt, OR_W);
    couple_vector(s);
    if (p->nu != 0.) {
        s->counters[0] = diffuse(s, s->u, s->u0, s->ut, p->nu, NULL, OR_U);
        s->counters[1] = diffuse(s, s->v, s->v0, s->vt, p->nu, NULL, OR_V);
        s->counters[2] = diffuse(s, s->w, s->w0, s->wt, p->nu, NULL, OR_W);
        couple_vector(s);
    }
    if (p->has_force) { /* ExplicitEulerSource vector add_source, ExplicitEulerSource.cpp:25-76 */
        if (has_dir(p->force_dir, 0)) or_add_source(D, s->u, s->fx, p->dt, OR_U);
        if (has_dir(p->force_dir, 1)) or_add_source(D, s->v, s->fy, p->dt, OR_V);
        if (has_dir(p->force_dir, 2)) or_add_source(D, s->w, s->fz, p->dt, OR_W);
        couple_vector(s);
    }
    or_divergence(D, s->rhs, s->ut, s->vt, s->wt);
    s->counters[4] = or_mg_pressure(s->mg, s->p_, s->rhs, t, p->dt);
    or_projection(D, s->u, s->v, s->w, s->ut, s->vt, s->wt, s->p_);
}

static void do_step_nstempturb(or_sim *s, double t) {
    const or_domain *D = s->D;
    const or_sim_params *p = &s->p;
    or_sl_advect(D, s->u, s->u0, s->u0, s->v0, s->w0, p->dt, OR_U);
    or_sl_advect(D, s->v, s->v0, s->u0, s->v0, s->w0, p->dt, OR_V);
    or_sl_advect(D, s->w, s->w0, s->u0, s->v0, s->w0, p->dt, OR_W);
    couple_vector(s);
    turbulence(s);
    s->counters[0] = diffuse(s, s->u, s->u0, s->ut, p->nu, s->nu_t, OR_U);
    s->counters[1] = diffuse(s, s->v, s->v0, s->vt, p->nu, s->nu_t, OR_V);
    s->counters[2] = diffuse(s, s->w, s->w0, s->wt, p->nu, s->nu_t, OR_W);
    couple_vector(s);
    if (p->has_force) {
        if (has_dir(p->force_dir, 0)) or_add_source(D, s->u, s->fx, p->dt, OR_U);
        if (has_dir(p->force_dir, 1)) or_add_source(D, s->v, s->fy, p->dt, OR_V);
        if (has_dir(p->force_dir, 2)) or_add_source(D, s->w, s->fz, p->dt, OR_W);
        couple_vector(s);
    }
    or_divergence(D, s->rhs, s->ut, s->vt, s->wt);
    s->counters[4] = or_mg_pressure(s->mg, s->p_, s->rhs, t, p->dt);
    or_projection(D, s->u, s->v, s->w, s->ut, s->vt, s->wt, s->p_);
    or_sl_advect(D, s->T, s->T0, s->u, s->v, s->w, p->dt, OR_T);
    couple_scalar(s);
    if (p->prandtl > 0) {
        const double rPr = 1. / p->prandtl;
        for (size_t i = 0; i < s->n; ++i) s->kappa_t[i] = s->nu_t[i] * rPr;
        s->counters[3] = diffuse(s, s->T, s->T0, s->Tt, p->kappa, s->kappa_t, OR_T);
        couple_scalar(s);
    } else if (p->kappa != 0.) {
        s->counters[3] = diffuse(s, s->T, s->T0, s->Tt, p->kappa, NULL, OR_T);
        couple_scalar(s);
    }
    if (p->solver == 10 && p->dissipation) { /* NSTempTurbConSolver.cpp:149-169 */
        or_dissipate(D, s->T, s->u, s->v, s->w, p->nu, p->dt, OR_T);
        couple_scalar(s);
    } else if (p->has_temp_source) {
        /* the scalar add_source ignores the direction list, ExplicitEulerSource.cpp:84-104 */
        or_add_source(D, s->T, s->S_T, p->dt, OR_T);
        couple_scalar(s);
    }
}

/* the concentration step, NSTempConSolver.cpp:204-233, NSTempTurbConSolver.cpp:241-286 */
static void do_step_con(or_sim *s) {
    const or_domain *D = s->D;
    const or_sim_params *p = &s->p;
    or_sl_advect(D, s->C, s->C0, s->u, s->v, s->w, p->dt, OR_RHO);
    couple_con(s);
    if (p->solver == 10 && p->con_turb) {
        const double rSc = 1. / p->schmidt;
        for (size_t i = 0; i < s->n; ++i) s->gamma_t[i] = s->nu_t[i] * rSc;
        diffuse(s, s->C, s->C0, s->Ct, p->gamma, s->gamma_t, OR_RHO);
        couple_con(s);
    } else if (p->gamma != 0.) {
        diffuse(s, s->C, s->C0, s->Ct, p->gamma, NULL, OR_RHO);
        couple_con(s);
    }
    /* NSTempTurbConSolver adds the source only for a turbulent concentration */
    if (p->solver == 10 ? p->con_turb : 1) {
        or_add_source(D, s->C, s->S_C, p->dt, OR_RHO);
        couple_con(s);
    }
}

static void vector_force(or_sim *s) {
    const or_sim_params *p = &s->p;
    if (p->has_force) {
        if (has_dir(p->force_dir, 0)) or_add_source(s->D, s->u, s->fx, p->dt, OR_U);
        if (has_dir(p->force_dir, 1)) or_add_source(s->D, s->v, s->fy, p->dt, OR_V);
        if (has_dir(p->force_dir, 2)) or_add_source(s->D, s->w, s->fz, p->dt, OR_W);
        couple_vector(s);
    }
}

/* NSTempSolver.cpp:69-198 */
static void do_step_nstemp(or_sim *s, double t) {
    const or_domain *D = s->D;
    const or_sim_params *p = &s->p;
    or_sl_advect(D, s->u, s->u0, s->u0, s->v0, s->w0, p->dt, OR_U);
    or_sl_advect(D, s->v, s->v0, s->u0, s->v0, s->w0, p->dt, OR_V);
    or_sl_advect(D, s->w, s->w0, s->u0, s->v0, s->w0, p->dt, OR_W);
    couple_vector(s);
    if (p->nu != 0.) {
        s->counters[0] = diffuse(s, s->u, s->u0, s->ut, p->nu, NULL, OR_U);
        s->counters[1] = diffuse(s, s->v, s->v0, s->vt, p->nu, NULL, OR_V);
        s->counters[2] = diffuse(s, s->w, s->w0, s->wt, p->nu, NULL, OR_W);
        couple_vector(s);
    }
    vector_force(s);
    or_divergence(D, s->rhs, s->ut, s->vt, s->wt);
    s->counters[4] = or_mg_pressure(s->mg, s->p_, s->rhs, t, p->dt);
    or_projection(D, s->u, s->v, s->w, s->ut, s->vt, s->wt, s->p_);
    or_sl_advect(D, s->T, s->T0, s->u, s->v, s->w, p->dt, OR_T);
    couple_scalar(s);
    if (p->kappa != 0.) {
        s->counters[3] = diffuse(s, s->T, s->T0, s->Tt, p->kappa, NULL, OR_T);
        couple_scalar(s);
    }
    if (p->dissipation) {
        or_dissipate(D, s->T, s->u, s->v, s->w, p->nu, p->dt, OR_T);
        couple_scalar(s);
    }
    if (p->has_temp_source) {
        or_add_source(D, s->T, s->S_T, p->dt, OR_T);
        couple_scalar(s);
    }
}

/* NSTurbSolver.cpp do_step */
static void do_step_nsturb(or_sim *s, double t) {
    const or_domain *D = s->D;
    const or_sim_params *p = &s->p;
    or_sl_advect(D, s->u, s->u0, s->u0, s->v0, s->w0, p->dt, OR_U);
    or_sl_advect(D, s->v, s->v0, s->u0, s->v0, s->w0, p->dt, OR_V);
    or_sl_advect(D, s->w, s->w0, s->u0, s->v0, s->w0, p->dt, OR_W);
    couple_vector(s);
    turbulence(s);
    s->counters[0] = diffuse(s, s->u, s->u0, s->ut, p->nu, s->nu_t, OR_U);
    s->counters[1] = diffuse(s, s->v, s->v0, s->vt, p->nu, s->nu_t, OR_V);
    s->counters[2] = diffuse(s, s->w, s->w0, s->wt, p->nu, s->nu_t, OR_W);
    couple_vector(s);
    vector_force(s);
    or_divergence(D, s->rhs, s->ut, s->vt, s->wt);
    s->counters[4] = or_mg_pressure(s->mg, s->p_, s->rhs, t, p->dt);
    or_projection(D, s->u, s->v, s->w, s->ut, s->vt, s->wt, s->p_);
}

/* DiffusionSolver, AdvectionDiffusionSolver, DiffusionTurbSolver, AdvectionSolver do_step */
static void do_step_transport(or_sim *s) {
    const or_domain *D = s->D;
    const or_sim_params *p = &s->p;
    if (p->solver == 8) {
        or_sl_advect(D, s->u, s->u0, s->lin[0], s->lin[1], s->lin[2], p->dt, OR_U);
        or_sl_advect(D, s->v, s->v0, s->lin[0], s->lin[1], s->lin[2], p->dt, OR_V);
        or_sl_advect(D, s->w, s->w0, s->lin[0], s->lin[1], s->lin[2], p->dt, OR_W);
        return;
    }
    if (p->solver == 6) {
        or_sl_advect(D, s->u, s->u0, s->u0, s->v0, s->w0, p->dt, OR_U);
        or_sl_advect(D, s->v, s->v0, s->u0, s->v0, s->w0, p->dt, OR_V);
        or_sl_advect(D, s->w, s->w0, s->u0, s->v0, s->w0, p->dt, OR_W);
        couple_vector(s);
        if (p->nu == 0.) return;
    }
    const double *ev = NULL;
    if (p->solver == 7) {
        turbulence(s);
        ev = s->nu_t;
    }
    s->counters[0] = diffuse(s, s->u, s->u0, s->ut, p->nu, ev, OR_U);
    s->counters[1] = diffuse(s, s->v, s->v0, s->vt, p->nu, ev, OR_V);
    s->counters[2] = diffuse(s, s->w, s->w0, s->wt, p->nu, ev, OR_W);
}

void or_sim_step(or_sim *s, double t) {
    memset(s->counters, 0, sizeof(s->counters));
    switch (s->p.solver) {
        case 0: do_step_ns(s, t); break;
        case 1: do_step_nstempturb(s, t); break;
        case 3: do_step_nstemp(s, t); break;
        case 4: do_step_nsturb(s, t); break;
        case 9:
            do_step_nstemp(s, t);
            do_step_con(s);
            break;
        case 10:
            do_step_nstempturb(s, t);
            do_step_con(s);
            break;
        default: do_step_transport(s); break;
    }
    update_sources(s, t);
    update_data(s);
}

double *or_sim_field(or_sim *s, const char *name) {
    if (!strcmp(name, "u")) return s->u;
    if (!strcmp(name, "v")) return s->v;
    if (!strcmp(name, "w")) return s->w;
    if (!strcmp(name, "p")) return s->p_;
    if (!strcmp(name, "T")) return s->T;
    if (!strcmp(name, "rhs")) return s->rhs;
    if (!strcmp(name, "nu_t")) return s->nu_t;
    if (!strcmp(name, "S_T")) return s->S_T;
    if (!strcmp(name, "f_x")) return s->fx;
    if (!strcmp(name, "f_y")) return s->fy;
    if (!strcmp(name, "f_z")) return s->fz;
    if (!strcmp(name, "T_ambient")) return s->Ta;
    if (!strcmp(name, "C") || !strcmp(name, "rho")) return s->C;
    if (!strcmp(name, "S_C")) return s->S_C;
    return NULL;
}

void or_sim_counters(const or_sim *s, int *out5) { memcpy(out5, s->counters, sizeof(s->counters)); }
