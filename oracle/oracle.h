/* CPU oracle for the ARTSS solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial C restatement of the reference
 * algorithm, used to check the HIP implementation in artss_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links, loads or falls back to it.
 *
 * Every function follows the reference operation by operation (same loop
 * order over the same index lists, same floating-point expression order, no
 * FMA contraction: built with -ffp-contract=off) and cites the reference
 * file:line it restates.  Field layout is the reference's: x fastest,
 * idx = i + Nx*j + Nx*Ny*k with one ghost layer (utility/GlobalMacrosTypes.h:70).
 *
 * Parity status of each function (how it is pinned) is listed in
 * tests/test_oracle_golden.py and in the module docstring there.
 */
#ifndef ARTSS_ORACLE_H
#define ARTSS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_OBST 32
#define OR_MAX_LEVELS 12

/* field types, utility/Mapping.h:48-51 */
enum { OR_RHO = 0, OR_U = 1, OR_V = 2, OR_W = 3, OR_P = 4, OR_T = 5, OR_NU = 6, OR_NFT = 7 };
/* boundary conditions, utility/Mapping.h:53-59 */
enum { OR_BC_NONE = -1, OR_NEUMANN = 0, OR_DIRICHLET = 1, OR_PERIODIC = 2 };
/* patches, utility/Mapping.h:39-47 */
enum { OR_LEFT = 0, OR_RIGHT = 1, OR_BOTTOM = 2, OR_TOP = 3, OR_FRONT = 4, OR_BACK = 5 };

typedef struct {
    int type;     /* OR_NEUMANN / OR_DIRICHLET / OR_PERIODIC */
    double value;
} or_bc;

typedef struct {
    double X1, X2, Y1, Y2, Z1, Z2; /* physical domain == computational domain */
    int nx, ny, nz;                /* inner cells at level 0 */
    int levels;                    /* multigrid levels (DomainData m_levels) */
    int n_obst;
    double obst[OR_MAX_OBST][6];   /* ox1, ox2, oy1, oy2, oz1, oz2 */
    /* boundary data: has[ft] = field type has any entry (BoundaryData::is_empty) */
    int dom_has[OR_NFT];
    or_bc dom[OR_NFT][6];
    int obst_has[OR_MAX_OBST][OR_NFT];
    or_bc obc[OR_MAX_OBST][OR_NFT][6];
} or_desc;

typedef struct or_domain or_domain;

/* ---- geometry (domain/DomainData.cpp, Domain.cpp, Obstacle.cpp, Multigrid.cpp) ---- */
or_domain *or_domain_create(const or_desc *d, char *err, size_t errlen);
void or_domain_free(or_domain *D);
void or_level_info(const or_domain *D, int level, size_t *NxNyNz_nxnynz_size /* [7] */, double *dxdydz /* [3] */);
/* which: 0 inner, 1 domain, 2 obstacle cells, 3+p boundary patch p, 10+p obstacle `id` face p */
size_t or_list(const or_domain *D, int which, int level, int id, const uint64_t **out);
void or_obstacle_box(const or_domain *D, int level, int id, size_t *start3, size_t *end3);

/* ---- boundary fill (domain/Multigrid.cpp:555-572) ---- */
void or_apply_bc(const or_domain *D, double *f, int ft, int level);

/* ---- operators at level 0 ---- */
void or_sl_advect(const or_domain *D, double *out, const double *in, const double *u, const double *v,
                  const double *w, double dt, int out_ft);
int or_jacobi_diffuse(const or_domain *D, double *out, double *in, const double *b, double Dc, const double *ev,
                      double dt, int max_iter, double tol_res, double wgt, int ft, double *res_out);
int or_cgs_diffuse(const or_domain *D, double *out, double *in, const double *b, double Dc, const double *ev,
                   double dt, int max_iter, double tol_res, double wgt, int ft, double *res_out);
void or_explicit_diffuse(const or_domain *D, double *out, const double *in, double Dc, const double *ev, double dt,
                         int ft);
void or_dissipate(const or_domain *D, double *out, const double *in_u, const double *in_v, const double *in_w,
                  double nu, double dt, int ft);
void or_divergence(const or_domain *D, double *out, const double *x, const double *y, const double *z);
void or_projection(const or_domain *D, double *ou, double *ov, double *ow, const double *iu, const double *iv,
                   const double *iw, const double *ip);
void or_smagorinsky(const or_domain *D, double *ev, const double *u, const double *v, const double *w, double cs,
                    int ev_ft);
/* DynamicSmagorinsky::calc_turbulent_viscosity (turbulence/DynamicSmagorinsky.cpp:172-393); nu the
 * molecular viscosity; ev written on the inner cells only */
void or_dyn_smagorinsky(const or_domain *D, double *ev, const double *u, const double *v, const double *w,
                        double nu);
void or_add_source(const or_domain *D, double *out, const double *s, double dt, int ft);
void or_buoyancy(const or_domain *D, double *out, const double *T, const double *Ta, double beta, double g);
/* GaussFunction::create_spatial_values (HRR, cp, x0,y0,z0, sx,sy,sz) */
void or_gauss_spatial(const or_domain *D, double *out, const double *par8);
/* L2 residual ||b - lap(x)|| over inner cells of `level` (VCycleMG.cpp:186-195 / 668-675) */
double or_residual_norm(const or_domain *D, const double *x, const double *b, int level);

/* ---- multigrid V-cycle pressure solver (pressure/VCycleMG.cpp) ---- */
typedef struct {
    int n_level, n_cycle, max_cycle, n_relax;
    double tol_res;
    int smoother;     /* 0 Jacobi, 1 ColoredGaussSeidel */
    int max_iter;     /* smoother max_iter */
    double s_tol_res; /* smoother tol_res */
    double w;         /* smoother weight */
} or_mg_params;
typedef struct or_mg or_mg;
or_mg *or_mg_create(const or_domain *D, const or_mg_params *p);
void or_mg_free(or_mg *mg);
/* returns cycles run; IPressure contract: out = p (in/out), b = rhs */
int or_mg_pressure(or_mg *mg, double *out, const double *b, double t, double dt);

/* ---- whole solver (solver/NSSolver.cpp, NSTempTurbSolver.cpp, TimeIntegration.cpp) ---- */
typedef struct {
    int solver;         /* 0 NSSolver, 1 NSTempTurbSolver, 3 NSTempSolver, 4 NSTurbSolver, 5 DiffusionSolver,
                           6 AdvectionDiffusionSolver, 7 DiffusionTurbSolver, 8 AdvectionSolver,
                           9 NSTempConSolver, 10 NSTempTurbConSolver */
    double dt, nu, kappa, beta, g;
    int diff_type;      /* 0 Jacobi, 1 CGS, 2 Explicit */
    int diff_max_iter;
    double diff_tol, diff_w;
    double cs;          /* ConstSmagorinsky */
    double prandtl;     /* temperature turbulence Pr_T; <=0: no turbulence */
    int has_force;      /* force_fct Buoyancy */
    int force_dir[3];   /* momentum source directions */
    double T_ambient;   /* buoyancy ambient_temperature_value */
    int has_temp_source;/* temp_fct Gauss */
    int temp_dir[3];
    double gauss[8];    /* HRR, cp, x0, y0, z0, sigma_x, sigma_y, sigma_z */
    double tau;
    or_mg_params mg;
    int dissipation;    /* NSTempSolver: temperature source dissipation */
    double u_lin[3];    /* AdvectionSolver: constant transport velocity */
    int turb_dynamic;   /* 1: DynamicSmagorinsky instead of ConstSmagorinsky */
    /* concentration solvers (9 NSTempConSolver, 10 NSTempTurbConSolver) */
    double gamma;       /* molecular diffusivity of the concentration */
    int con_turb;       /* NSTempTurbConSolver: <turbulence include="Yes"> */
    double schmidt;     /* Sc_T */
    int con_dir[3];
    int con_fct;        /* 1 Gauss, 2 Cube */
    double con_gauss[8];
    double con_tau;
    double con_cube[7]; /* x, y, z start, x, y, z end, value */
} or_sim_params;
typedef struct or_sim or_sim;
/* u_ic ... p_ic: full arrays whose inner cells carry the initial condition (may be NULL) */
or_sim *or_sim_create(const or_domain *D, const or_sim_params *p, const double *u_ic, const double *v_ic,
                      const double *w_ic, const double *T_ic, const double *p_ic);
void or_sim_free(or_sim *s);
/* one TimeIntegration iteration: do_step(t), update_sources(t), update_data */
void or_sim_step(or_sim *s, double t);
/* name: u v w p T rhs nu_t S_T f_x f_y f_z ; returns pointer to level-0 array */
double *or_sim_field(or_sim *s, const char *name);
/* iteration counters of the last step: [ju, jv, jw, jT, mg_cycles] */
void or_sim_counters(const or_sim *s, int *out5);

#ifdef __cplusplus
}
#endif
#endif
