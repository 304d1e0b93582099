/* Oracle geometric multigrid pressure solver.  TEST INFRASTRUCTURE ONLY.
 * Restates pressure/VCycleMG.cpp, including the Jacobi pointer-swap
 * semantics (:685-798), the dead store at :254 and the effective coarsest
 * level n_level-1 (:273-279). */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double *data;
    int level;
} fld;

struct or_mg {
    const or_domain *D;
    or_mg_params p;
    fld res0[OR_MAX_LEVELS], res1[OR_MAX_LEVELS], err0[OR_MAX_LEVELS], err1[OR_MAX_LEVELS], tmp[OR_MAX_LEVELS];
    ivec odd[OR_MAX_LEVELS], even[OR_MAX_LEVELS];
};

static const double DSIGN = -1; /* VCycleMG.h:53 */

static fld fld_new(const or_domain *D, int level) {
    fld f;
    f.level = level;
    f.data = (double *)calloc(D->size[level], sizeof(double));
    return f;
}
static void fcopy(const or_domain *D, fld *dst, const fld *src) {
    memcpy(dst->data, src->data, D->size[dst->level] * sizeof(double));
}
static void fswap(fld *a, fld *b) {
    double *t = a->data;
    a->data = b->data;
    b->data = t;
}

or_mg *or_mg_create(const or_domain *D, const or_mg_params *p) {
    or_mg *mg = (or_mg *)calloc(1, sizeof(or_mg));
    mg->D = D;
    mg->p = *p;
    const int L = p->n_level;
    mg->res1[0] = fld_new(D, 0);
    mg->err1[0] = fld_new(D, 0);
    mg->tmp[0] = fld_new(D, 0);
    for (int l = 0; l < L; l++) {
        mg->res0[l] = fld_new(D, l);
        mg->res1[l + 1] = fld_new(D, l + 1);
        mg->err0[l] = fld_new(D, l);
        mg->err1[l + 1] = fld_new(D, l + 1);
        mg->tmp[l + 1] = fld_new(D, l + 1);
    }
    if (p->smoother == 1) { /* ColoredGaussSeidelDiffuse::create_red_black_lists per level, VCycleMG.cpp:38-42 */
        for (int l = 0; l <= L; l++) {
            const ivec *in = &D->inner[l];
            const size_t Nx = D->N[l][0], Ny = D->N[l][1];
            mg->odd[l].d = (uint64_t *)malloc((in->n + 1) * sizeof(uint64_t));
            mg->even[l].d = (uint64_t *)malloc((in->n + 1) * sizeof(uint64_t));
            for (size_t c = 0; c < in->n; c++) {
                size_t idx = in->d[c];
                size_t k = idx / (Nx * Ny), j = (idx - k * Nx * Ny) / Nx, i = idx - k * Nx * Ny - j * Nx;
                if ((i + j + k) % 2 == 0) mg->even[l].d[mg->even[l].n++] = idx;
                else mg->odd[l].d[mg->odd[l].n++] = idx;
            }
        }
    }
    return mg;
}

void or_mg_free(or_mg *mg) {
    if (!mg) return;
    const int L = mg->p.n_level;
    free(mg->res1[0].data);
    free(mg->err1[0].data);
    free(mg->tmp[0].data);
    for (int l = 0; l < L; l++) {
        free(mg->res0[l].data);
        free(mg->res1[l + 1].data);
        free(mg->err0[l].data);
        free(mg->err1[l + 1].data);
        free(mg->tmp[l + 1].data);
    }
    for (int l = 0; l <= L; l++) {
        free(mg->odd[l].d);
        free(mg->even[l].d);
    }
    free(mg);
}

static void bc(or_mg *mg, fld *f) { or_apply_bc(mg->D, f->data, OR_P, f->level); }

/* JacobiDiffuse::JacobiStep multigrid form, JacobiDiffuse.cpp:274-305 */
static void jacobi_step_mg(or_mg *mg, int level, fld *out, const fld *in, const fld *b, double ax, double ay,
                           double az, double beta, double w) {
    const or_domain *D = mg->D;
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    const size_t nj = Nx, nk = Nx * Ny;
    const ivec *inner = &D->inner[level];
    double *o = out->data;
    const double *x = in->data, *bb = b->data;
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        double out_h = beta * (DSIGN * bb[i] + ax * (x[i + 1] + x[i - 1]) + ay * (x[i + nj] + x[i - nj]) +
                               az * (x[i + nk] + x[i - nk]));
        o[i] = (1 - w) * x[i] + w * out_h;
    }
}

/* ColoredGaussSeidelDiffuse::colored_gauss_seidel_step, :190-224 */
static void cgs_step_mg(or_mg *mg, int level, fld *out, const fld *b, double ax, double ay, double az, double beta,
                        double w) {
    const or_domain *D = mg->D;
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    const size_t nj = Nx, nk = Nx * Ny;
    double *o = out->data;
    const double *bb = b->data;
    const ivec *col[2] = {&mg->even[level], &mg->odd[level]};
    for (int c = 0; c < 2; c++) {
        for (size_t n = 0; n < col[c]->n; n++) {
            const size_t index = col[c]->d[n];
            double ox = o[index + 1], ox2 = o[index - 1], oy = o[index + nj], oy2 = o[index - nj];
            double oz = o[index + nk], oz2 = o[index - nk], db = bb[index], r_out = o[index];
            double out_h = beta * (DSIGN * db + ax * (ox + ox2) + ay * (oy + oy2) + az * (oz + oz2));
            o[index] = (1 - w) * r_out + w * out_h;
        }
    }
}

static void level_coeffs(or_mg *mg, int level, double *rdx2, double *rdy2, double *rdz2, double *beta) {
    const double dx = mg->D->h[level][0], dy = mg->D->h[level][1], dz = mg->D->h[level][2];
    *rdx2 = 1. / (dx * dx);
    *rdy2 = 1. / (dy * dy);
    *rdz2 = 1. / (dz * dz);
    const double rb = 2. * (*rdx2 + *rdy2 + *rdz2);
    *beta = 1. / rb;
}

static double level_residual(or_mg *mg, const fld *out, const fld *b, int level) {
    return or_residual_norm(mg->D, out->data, b->data, level);
}

/* VCycleMG::call_smooth_jacobi :685-723 / call_smooth_colored_gauss_seidel :583-615 */
static void smooth_fn(or_mg *mg, fld *out, fld *tmp, const fld *b, int level) {
    double ax, ay, az, beta;
    bc(mg, out);
    level_coeffs(mg, level, &ax, &ay, &az, &beta);
    if (mg->p.smoother == 0) {
        fcopy(mg->D, tmp, out);
        int it = 0;
        for (int i = 0; i < mg->p.n_relax; i++) {
            jacobi_step_mg(mg, level, out, tmp, b, ax, ay, az, beta, mg->p.w);
            bc(mg, out);
            fswap(tmp, out);
            it++;
        }
        if (it % 2 != 0) fswap(tmp, out);
    } else {
        for (int i = 0; i < mg->p.n_relax; i++) {
            cgs_step_mg(mg, level, out, b, ax, ay, az, beta, mg->p.w);
            bc(mg, out);
        }
    }
}

/* VCycleMG::call_solve_jacobi :725-798 / call_solve_colored_gauss_seidel :617-683 */
static void solve_fn(or_mg *mg, fld *out, fld *tmp, const fld *b, int level) {
    double ax, ay, az, beta;
    bc(mg, out);
    level_coeffs(mg, level, &ax, &ay, &az, &beta);
    int it = 0;
    double res = 10000.;
    if (mg->p.smoother == 0) {
        fcopy(mg->D, tmp, out);
        while (res > mg->p.s_tol_res && it < mg->p.max_iter) {
            jacobi_step_mg(mg, level, out, tmp, b, ax, ay, az, beta, mg->p.w);
            bc(mg, out);
            res = level_residual(mg, out, b, level);
            fswap(tmp, out);
            it++;
        }
        if (it % 2 != 0) fswap(tmp, out);
    } else {
        while (res > mg->p.s_tol_res && it < mg->p.max_iter) {
            cgs_step_mg(mg, level, out, b, ax, ay, az, beta, mg->p.w);
            bc(mg, out);
            res = level_residual(mg, out, b, level);
            it++;
        }
    }
}

/* VCycleMG::Smooth :299-308 */
static void Smooth(or_mg *mg, fld *out, fld *tmp, const fld *b, int level) {
    bc(mg, out);
    smooth_fn(mg, out, tmp, b, level);
}

/* VCycleMG::Solve :562-581 */
static void Solve(or_mg *mg, fld *out, fld *tmp, const fld *b, int level) {
    const size_t Nx = mg->D->N[out->level][0], Ny = mg->D->N[out->level][1];
    if (Nx <= 4 && Ny <= 4) {
        Smooth(mg, out, tmp, b, level);
        return;
    }
    solve_fn(mg, out, tmp, b, level);
}

/* VCycleMG::Residuum :319-360 */
static void Residuum(or_mg *mg, fld *out, const fld *in, const fld *b, int level) {
    const or_domain *D = mg->D;
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    const size_t nj = Nx, nk = Nx * Ny;
    double rdx2, rdy2, rdz2, beta;
    level_coeffs(mg, level, &rdx2, &rdy2, &rdz2, &beta);
    const ivec *inner = &D->inner[level];
    const double *x = in->data;
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        double weighted = (rdx2 * (x[i - 1] - 2 * x[i] + x[i + 1]) + rdy2 * (x[i - nj] - 2 * x[i] + x[i + nj]) +
                           rdz2 * (x[i - nk] - 2 * x[i] + x[i + nk]));
        out->data[i] = b->data[i] - weighted;
    }
}

/* VCycleMG::Restrict :370-427 (coarse inner list of level+1) */
static void Restrict(or_mg *mg, fld *out, const fld *in, int level) {
    const or_domain *D = mg->D;
    const size_t Nxc = D->N[out->level][0], Nyc = D->N[out->level][1];
    const size_t Nxf = D->N[in->level][0], Nyf = D->N[in->level][1];
    const ivec *inner = &D->inner[level + 1];
    const double *x = in->data;
#define F(a, b, c) x[(a) + Nxf * (b) + Nxf * Nyf * (c)]
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t idx = inner->d[l];
        const size_t k = idx / (Nxc * Nyc), j = (idx - k * Nxc * Nyc) / Nxc, i = idx - k * Nxc * Nyc - j * Nxc;
        out->data[idx] = 0.125 * (F(2 * i - 1, 2 * j - 1, 2 * k - 1) + F(2 * i, 2 * j - 1, 2 * k - 1) +
                                  F(2 * i - 1, 2 * j, 2 * k - 1) + F(2 * i, 2 * j, 2 * k - 1) +
                                  F(2 * i - 1, 2 * j - 1, 2 * k) + F(2 * i, 2 * j - 1, 2 * k) +
                                  F(2 * i - 1, 2 * j, 2 * k) + F(2 * i, 2 * j, 2 * k));
    }
#undef F
}

/* VCycleMG::Prolongate :437-551 (coarse inner list of `level`) */
static void Prolongate(or_mg *mg, fld *out, const fld *in, int level) {
    const or_domain *D = mg->D;
    const size_t Nxf = D->N[out->level][0], Nyf = D->N[out->level][1];
    const size_t Nxc = D->N[in->level][0], Nyc = D->N[in->level][1];
    const size_t ni = 1, nj = Nxc, nk = Nxc * Nyc;
    const ivec *inner = &D->inner[level];
    const double *c = in->data;
    double *o = out->data;
#define IXF(a, b, d) ((a) + Nxf * (b) + Nxf * Nyf * (d))
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t idx = inner->d[l];
        const size_t k = idx / (Nxc * Nyc), j = (idx - k * Nxc * Nyc) / Nxc, i = idx - k * Nxc * Nyc - j * Nxc;
        o[IXF(2 * i, 2 * j, 2 * k)] =
            0.015625 * (27 * c[idx] + 9 * c[idx + ni] + 9 * c[idx + nj] + 9 * c[idx + nk] + 3 * c[idx + ni + nj] +
                        3 * c[idx + ni + nk] + 3 * c[idx + nj + nk] + c[idx + ni + nj + nk]);
        o[IXF(2 * i, 2 * j, 2 * k - 1)] =
            0.015625 * (27 * c[idx] + 9 * c[idx + ni] + 9 * c[idx + nj] + 9 * c[idx - nk] + 3 * c[idx + ni + nj] +
                        3 * c[idx + ni - nk] + 3 * c[idx + nj - nk] + c[idx + ni + nj - nk]);
        o[IXF(2 * i, 2 * j - 1, 2 * k)] =
            0.015625 * (27 * c[idx] + 9 * c[idx + ni] + 9 * c[idx - nj] + 9 * c[idx + nk] + 3 * c[idx + ni - nj] +
                        3 * c[idx + ni + nk] + 3 * c[idx - nj + nk] + c[idx + ni - nj + nk]);
        o[IXF(2 * i, 2 * j - 1, 2 * k - 1)] =
            0.015625 * (27 * c[idx] + 9 * c[idx + ni] + 9 * c[idx - nj] + 9 * c[idx - nk] + 3 * c[idx + ni - nj] +
                        3 * c[idx + ni - nk] + 3 * c[idx - nj - nk] + c[idx + ni - nj - nk]);
        o[IXF(2 * i - 1, 2 * j, 2 * k)] =
            0.015625 * (27 * c[idx] + 9 * c[idx - ni] + 9 * c[idx + nj] + 9 * c[idx + nk] + 3 * c[idx - ni + nj] +
                        3 * c[idx - ni + nk] + 3 * c[idx + nj + nk] + c[idx - ni + nj + nk]);
        o[IXF(2 * i - 1, 2 * j, 2 * k - 1)] =
            0.015625 * (27 * c[idx] + 9 * c[idx - ni] + 9 * c[idx + nj] + 9 * c[idx - nk] + 3 * c[idx - ni + nj] +
                        3 * c[idx - ni - nk] + 3 * c[idx + nj - nk] + c[idx - ni + nj - nk]);
        o[IXF(2 * i - 1, 2 * j - 1, 2 * k)] =
            0.015625 * (27 * c[idx] + 9 * c[idx - ni] + 9 * c[idx - nj] + 9 * c[idx + nk] + 3 * c[idx - ni - nj] +
                        3 * c[idx - ni + nk] + 3 * c[idx - nj + nk] + c[idx - ni - nj + nk]);
        o[IXF(2 * i - 1, 2 * j - 1, 2 * k - 1)] =
            0.015625 * (27 * c[idx] + 9 * c[idx - ni] + 9 * c[idx - nj] + 9 * c[idx - nk] + 3 * c[idx - ni - nj] +
                        3 * c[idx - ni - nk] + 3 * c[idx - nj - nk] + c[idx - ni - nj - nk]);
    }
#undef IXF
}

/* VCycleMG::VCycleMultigrid :215-288 */
static void vcycle(or_mg *mg, fld *out) {
    const or_domain *D = mg->D;
    const int L = mg->p.n_level;
    if (L == 0) {
        Solve(mg, out, &mg->tmp[0], &mg->res1[0], 0);
        return;
    }
    fcopy(D, &mg->err1[0], out);
    for (int level = 0; level < L; ++level) {
        Smooth(mg, &mg->err1[level], &mg->tmp[level], &mg->res1[level], level);
        Residuum(mg, &mg->res0[level], &mg->err1[level], &mg->res1[level], level);
        bc(mg, &mg->res0[level]);
        Restrict(mg, &mg->res1[level + 1], &mg->res0[level], level);
        bc(mg, &mg->res1[level + 1]);
        memset(mg->err1[level + 1].data, 0, D->size[level + 1] * sizeof(double));
    }
    fcopy(D, out, &mg->err0[0]); /* dead store, :254 */
    for (int level = L; level > 0; --level) {
        Prolongate(mg, &mg->err0[level - 1], &mg->err1[level], level);
        bc(mg, &mg->err0[level - 1]);
        double *a = mg->err1[level - 1].data;
        const double *e = mg->err0[level - 1].data;
        for (size_t i = 0; i < D->size[level - 1]; i++) a[i] += e[i];
        if (level == L) Solve(mg, &mg->err1[level - 1], &mg->tmp[level - 1], &mg->res1[level - 1], level - 1);
        else Smooth(mg, &mg->err1[level - 1], &mg->tmp[level - 1], &mg->res1[level - 1], level - 1);
    }
    fcopy(D, out, &mg->err1[0]);
    bc(mg, out);
}

/* VCycleMG::pressure :142-207 */
int or_mg_pressure(or_mg *mg, double *out_data, const double *b_data, double t, double dt) {
    const or_domain *D = mg->D;
    fld out = {out_data, 0};
    /* UpdateInput :125-132 */
    memcpy(mg->err1[0].data, out_data, D->size[0] * sizeof(double));
    memcpy(mg->tmp[0].data, out_data, D->size[0] * sizeof(double));
    memcpy(mg->res1[0].data, b_data, D->size[0] * sizeof(double));
    const size_t Nt = (size_t)round(t / dt);
    int act = 0;
    if (Nt == 1) {
        size_t set_relax = (size_t)mg->p.n_relax;
        double r = 10000.;
        while (r > mg->p.tol_res && act < mg->p.max_cycle && set_relax < (size_t)mg->p.max_iter) {
            for (int i = 0; i < mg->p.n_cycle; i++) {
                vcycle(mg, &out);
                act++;
            }
            r = or_residual_norm(D, out_data, b_data, 0);
            set_relax += (size_t)mg->p.n_relax;
        }
    } else {
        for (int i = 0; i < mg->p.n_cycle; i++) {
            vcycle(mg, &out);
            act++;
        }
    }
    return act;
}
