/* Oracle operators at level 0 plus the boundary fill at any level.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function restates the cited
 * reference function with its loop order and expression order. */
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- boundary fill ---------------- */

/* DomainBoundary.cpp:25-49 and ObstacleBoundary.cpp:26-51: the same list kernel */
static void bc_list(double *f, const ivec *l, int sign_ref, size_t ref, double value, int sign) {
    const double s = (double)sign;
    for (size_t j = 0; j < l->n; j++) {
        const size_t idx = l->d[j];
        const size_t nb = sign_ref > 0 ? idx + ref : idx - ref;
        f[idx] = s * f[nb] + value;
    }
}

/* ObstacleBoundary::apply_boundary_condition, ObstacleBoundary.cpp:62-213 */
static void bc_obstacle(const or_domain *D, double *f, int ft, int level, int id) {
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    for (int p = 0; p < 6; p++) {
        const ivec *l = &D->obst[level][id].face[p];
        if (l->n == 0) continue;
        const or_bc bc = D->desc.obc[id][ft][p];
        size_t ref = 0;
        if (p == OR_FRONT || p == OR_BACK) ref = Nx * Ny;
        else if (p == OR_BOTTOM || p == OR_TOP) ref = Nx;
        else ref = 1;
        const int sign_ref = (p == OR_FRONT || p == OR_BOTTOM || p == OR_LEFT) ? -1 : 1;
        double value = bc.value;
        if (bc.type == OR_DIRICHLET) {
            if (level > 0) value = 0;
            bc_list(f, l, sign_ref, ref, value * 2, -1);
        } else if (bc.type == OR_NEUMANN) {
            if (level > 0) value = 0;
            if (p == OR_FRONT || p == OR_BACK) value *= D->h[level][2];
            else if (p == OR_BOTTOM || p == OR_TOP) value *= D->h[level][1];
            else value *= D->h[level][0];
            bc_list(f, l, sign_ref, ref, -value, 1);
        }
        /* periodic: not applied to obstacles (ObstacleBoundary.cpp:198-202) */
    }
}

/* DomainBoundary::apply_boundary_condition, DomainBoundary.cpp:59-239 */
static void bc_domain(const or_domain *D, double *f, int ft, int level) {
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    for (int p = 0; p < 6; p++) {
        const ivec *l = &D->bnd[level][p];
        if (l->n == 0) continue;
        const or_bc bc = D->desc.dom[ft][p];
        const int sign_ref = (p == OR_BACK || p == OR_TOP || p == OR_RIGHT) ? -1 : 1;
        double value = 0;
        size_t ref = 0;
        if (bc.type == OR_DIRICHLET) {
            if (level == 0) value = bc.value;
            if (p == OR_FRONT || p == OR_BACK) ref = Nx * Ny;
            else if (p == OR_BOTTOM || p == OR_TOP) ref = Nx;
            else ref = 1;
            bc_list(f, l, sign_ref, ref, value * 2, -1);
        } else if (bc.type == OR_NEUMANN) {
            if (level == 0) value = bc.value;
            if (p == OR_FRONT || p == OR_BACK) { value *= D->h[level][2]; ref = Nx * Ny; }
            else if (p == OR_BOTTOM || p == OR_TOP) { value *= D->h[level][1]; ref = Nx; }
            else { value *= D->h[level][2]; ref = 1; } /* dz for LEFT/RIGHT: DomainBoundary.cpp:120 */
            bc_list(f, l, sign_ref, ref, value, 1);
        } else if (bc.type == OR_PERIODIC) {
            if (p == OR_FRONT || p == OR_BACK) ref = Nx * Ny * D->n[level][2];
            else if (p == OR_BOTTOM || p == OR_TOP) ref = Nx * D->n[level][1];
            else ref = D->n[level][0];
            bc_list(f, l, sign_ref, ref, 0, 1);
        }
    }
}

/* Multigrid::apply_boundary_condition, Multigrid.cpp:555-572: obstacles, then domain */
void or_apply_bc(const or_domain *D, double *f, int ft, int level) {
    for (int id = 0; id < D->n_obst; id++)
        if (D->desc.obst_has[id][ft]) bc_obstacle(D, f, ft, level, id);
    if (D->desc.dom_has[ft]) bc_domain(D, f, ft, level);
}

/* ---------------- SL advection, advection/SLAdvect.cpp:34-170 ---------------- */
void or_sl_advect(const or_domain *D, double *out, const double *in, const double *u, const double *v,
                  const double *w, double dt, int out_ft) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const double dx = D->h[0][0], dy = D->h[0][1], dz = D->h[0][2];
    const double dtx = dt / dx, dty = dt / dy, dtz = dt / dz;
    const double eps = 1e-6;
    const long i_start = 1, j_start = 1, k_start = 1;
    const long i_end = (long)D->n[0][0], j_end = (long)D->n[0][1], k_end = (long)D->n[0][2];
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t idx = inner->d[l];
        long k = (long)(idx / (Nx * Ny));
        long j = (long)((idx - (size_t)k * Nx * Ny) / Nx);
        long i = (long)(idx - (size_t)k * Nx * Ny - (size_t)j * Nx);
        double Ci = dtx * u[idx], Cj = dty * v[idx], Ck = dtz * w[idx];
        long i0 = i, i1 = i, j0 = j, j1 = j, k0 = k, k1 = k;
        double r = 1, s = 1, t = 1;
        if (Ci > eps) {
            i0 = i - (long)Ci; if (i0 < i_start) i0 = i_start;
            i1 = i0 - 1;
            r = fmod(Ci, 1);
        } else if (Ci < -eps) {
            i0 = i - (long)Ci; if (i0 > i_end) i0 = i_end;
            i1 = i0 + 1;
            r = fmod(-Ci, 1);
        }
        if (Cj > eps) {
            j0 = j - (long)Cj; if (j0 < j_start) j0 = j_start;
            j1 = j0 - 1;
            s = fmod(Cj, 1);
        } else if (Cj < -eps) {
            j0 = j - (long)Cj; if (j0 > j_end) j0 = j_end;
            j1 = j0 + 1;
            s = fmod(-Cj, 1);
        }
        if (Ck > eps) {
            k0 = k - (long)Ck; if (k0 < k_start) k0 = k_start;
            k1 = k0 - 1;
            t = fmod(Ck, 1);
        } else if (Ck < -eps) {
            k0 = k - (long)Ck; if (k0 > k_end) k0 = k_end;
            k1 = k0 + 1;
            t = fmod(-Ck, 1);
        }
#define AT(a, b, c) in[(size_t)(a) + Nx * (size_t)(b) + Nx * Ny * (size_t)(c)]
        double d000 = AT(i0, j0, k0), d100 = AT(i1, j0, k0), d010 = AT(i0, j1, k0), d110 = AT(i1, j1, k0);
        double d001 = AT(i0, j0, k1), d101 = AT(i1, j0, k1), d011 = AT(i0, j1, k1), d111 = AT(i1, j1, k1);
#undef AT
        double r100 = d000 + r * (d100 - d000);
        double r110 = d010 + r * (d110 - d010);
        double r101 = d001 + r * (d101 - d001);
        double r111 = d011 + r * (d111 - d011);
        double s110 = r100 + s * (r110 - r100);
        double s111 = r101 + s * (r111 - r101);
        out[idx] = s110 + t * (s111 - s110);
    }
    or_apply_bc(D, out, out_ft, 0);
}

/* ---------------- Jacobi diffusion, diffusion/JacobiDiffuse.cpp ---------------- */
int or_jacobi_diffuse(const or_domain *D, double *out, double *in, const double *b, double Dc, const double *ev,
                      double dt, int max_iter, double tol_res, double wgt, int ft, double *res_out) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const size_t nj = Nx, nk = Nx * Ny;
    const double rdx = 1. / D->h[0][0], rdy = 1. / D->h[0][1], rdz = 1. / D->h[0][2];
    const double dsign = 1;
    const ivec *inner = &D->inner[0];
    const ivec *dom = &D->domain[0];
    /* plain version constants, JacobiDiffuse.cpp:54-58 */
    const double ax = Dc * dt * rdx * rdx, ay = Dc * dt * rdy * rdy, az = Dc * dt * rdz * rdz;
    const double rb_c = (1. + 2. * (ax + ay + az));
    int it = 0;
    double res = 1.;
    while (res > tol_res && it < max_iter) {
        for (size_t l = 0; l < inner->n; ++l) {
            const size_t i = inner->d[l];
            double out_h;
            if (!ev) { /* JacobiStep level-0 form, :241-249 (divides by reciprocal_beta) */
                out_h = (dsign * b[i] + ax * (in[i + 1] + in[i - 1]) + ay * (in[i + nj] + in[i - nj]) +
                         az * (in[i + nk] + in[i - nk])) / rb_c;
            } else { /* turbulent JacobiStep, :345-362 */
                double aX = (Dc + ev[i]) * dt * rdx * rdx;
                double aY = (Dc + ev[i]) * dt * rdy * rdy;
                double aZ = (Dc + ev[i]) * dt * rdz * rdz;
                double rb = (1. + 2. * (aX + aY + aZ));
                double bb = 1. / rb;
                double d_aX = aX * (in[i + 1] + in[i - 1]);
                double d_aY = aY * (in[i + nj] + in[i - nj]);
                double d_aZ = aZ * (in[i + nk] + in[i - nk]);
                out_h = bb * (dsign * b[i] + d_aX + d_aY + d_aZ);
            }
            out[i] = (1 - wgt) * in[i] + wgt * out_h;
        }
        or_apply_bc(D, out, ft, 0);
        double sum = 0.;
        for (size_t l = 0; l < inner->n; ++l) {
            const size_t i = inner->d[l];
            double rb = rb_c;
            if (ev) {
                double aX = (Dc + ev[i]) * dt * rdx * rdx;
                double aY = (Dc + ev[i]) * dt * rdy * rdy;
                double aZ = (Dc + ev[i]) * dt * rdz * rdz;
                rb = (1. + 2. * (aX + aY + aZ));
            }
            double r = rb * (out[i] - in[i]);
            sum += r * r;
        }
        res = sqrt(sum);
        it++;
        for (size_t l = 0; l < dom->n; ++l) in[dom->d[l]] = out[dom->d[l]];
    }
    if (it % 2 != 0)
        for (size_t l = 0; l < dom->n; ++l) in[dom->d[l]] = out[dom->d[l]];
    if (res_out) *res_out = res;
    return it;
}

/* ---------------- viscous dissipation, interfaces/ISource.cpp:63-121 ---------------- */
void or_dissipate(const or_domain *D, double *out, const double *in_u, const double *in_v, const double *in_w,
                  double nu, double dt, int ft) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const size_t ni = 1, nj = Nx, nk = Nx * Ny;
    const double rdx = 1. / D->h[0][0], rdy = 1. / D->h[0][1], rdz = 1. / D->h[0][2];
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        double out_h = nu * (2 * (0.5 * rdx * (in_u[i + ni] - in_u[i - ni])) * (0.5 * rdx * (in_u[i + ni] - in_u[i - ni])) +
                             2 * (0.5 * rdy * (in_v[i + nj] - in_v[i - nj])) * (0.5 * rdy * (in_v[i + nj] - in_v[i - nj])) +
                             2 * (0.5 * rdz * (in_w[i + nk] - in_w[i - nk])) * (0.5 * rdz * (in_w[i + nk] - in_w[i - nk])) +
                             ((0.5 * rdx * (in_v[i + ni] - in_v[i - ni])) + (0.5 * rdy * (in_u[i + nj] - in_u[i - nj]))) *
                                 ((0.5 * rdx * (in_v[i + ni] - in_v[i - ni])) + (0.5 * rdy * (in_u[i + nj] - in_u[i - nj]))) +
                             ((0.5 * rdy * (in_w[i + nj] - in_w[i - nj])) + (0.5 * rdz * (in_v[i + nk] - in_v[i - nk]))) *
                                 ((0.5 * rdy * (in_w[i + nj] - in_w[i - nj])) + (0.5 * rdz * (in_v[i + nk] - in_v[i - nk]))) +
                             ((0.5 * rdz * (in_u[i + nk] - in_u[i - nk])) + (0.5 * rdx * (in_w[i + ni] - in_w[i - ni]))) *
                                 ((0.5 * rdz * (in_u[i + nk] - in_u[i - nk])) + (0.5 * rdx * (in_w[i + ni] - in_w[i - ni]))));
        out[i] += dt * out_h;
    }
    or_apply_bc(D, out, ft, 0);
}

/* ---------------- explicit diffusion, diffusion/ExplicitDiffuse.cpp ---------------- */
/* one explicit Euler step on the inner list (:49-78 plain, :81-121 turbulent:
   the turbulent form has no 1/h^2 factors and subtracts the upper flux from
   the lower one, both as in the reference), then the boundary fill (:27, :44) */
void or_explicit_diffuse(const or_domain *D, double *out, const double *in, double Dc, const double *ev, double dt,
                         int ft) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const size_t ni = 1, nj = Nx, nk = Nx * Ny;
    const ivec *inner = &D->inner[0];
    if (!ev) {
        const double rdx = Dc / (D->h[0][0] * D->h[0][0]);
        const double rdy = Dc / (D->h[0][1] * D->h[0][1]);
        const double rdz = Dc / (D->h[0][2] * D->h[0][2]);
        for (size_t l = 0; l < inner->n; ++l) {
            const size_t i = inner->d[l];
            out[i] = in[i] + dt * ((in[i + ni] - 2 * in[i] + in[i - ni]) * rdx +
                                   (in[i + nj] - 2 * in[i] + in[i - nj]) * rdy +
                                   (in[i + nk] - 2 * in[i] + in[i - nk]) * rdz);
        }
    } else {
        for (size_t l = 0; l < inner->n; ++l) {
            const size_t i = inner->d[l];
            double nu_x1 = 0.5 * (ev[i + ni] + ev[i]) + Dc;
            double nu_x2 = 0.5 * (ev[i - ni] + ev[i]) + Dc;
            double nu_y1 = 0.5 * (ev[i + nj] + ev[i]) + Dc;
            double nu_y2 = 0.5 * (ev[i - nj] + ev[i]) + Dc;
            double nu_z1 = 0.5 * (ev[i + nk] + ev[i]) + Dc;
            double nu_z2 = 0.5 * (ev[i - nk] + ev[i]) + Dc;
            double di0 = (in[i] - in[i - ni]);
            double di1 = (in[i + ni] - in[i]);
            double dj0 = (in[i] - in[i - nj]);
            double dj1 = (in[i + nj] - in[i]);
            double dk0 = (in[i] - in[i - nk]);
            double dk1 = (in[i + nk] - in[i]);
            out[i] = in[i] + dt * ((nu_x1 * di0 - nu_x2 * di1) + (nu_y1 * dj0 - nu_y2 * dj1) +
                                   (nu_z1 * dk0 - nu_z2 * dk1));
        }
    }
    or_apply_bc(D, out, ft, 0);
}

/* ---------------- colored Gauss-Seidel, diffusion/ColoredGaussSeidelDiffuse.cpp ---------------- */
static void cgs_stencil(size_t index, double *out, const double *b, double ax, double ay, double az, double dsign,
                        double beta, double w, size_t Nx, size_t Ny) {
    const size_t nj = Nx, nk = Nx * Ny;
    double ox = out[index + 1], ox2 = out[index - 1];
    double oy = out[index + nj], oy2 = out[index - nj];
    double oz = out[index + nk], oz2 = out[index - nk];
    double db = b[index], r_out = out[index];
    double out_h = beta * (dsign * db + ax * (ox + ox2) + ay * (oy + oy2) + az * (oz + oz2));
    out[index] = (1 - w) * r_out + w * out_h;
}

int or_cgs_diffuse(const or_domain *D, double *out, double *in, const double *b, double Dc, const double *ev,
                   double dt, int max_iter, double tol_res, double wgt, int ft, double *res_out) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const size_t size = D->size[0];
    const double rdx = 1. / D->h[0][0], rdy = 1. / D->h[0][1], rdz = 1. / D->h[0][2];
    const double dsign = 1;
    const ivec *inner = &D->inner[0];
    /* create_red_black_lists(0), :343-372 */
    ivec even = {0}, odd = {0};
    even.d = (uint64_t *)malloc((inner->n + 1) * sizeof(uint64_t));
    odd.d = (uint64_t *)malloc((inner->n + 1) * sizeof(uint64_t));
    for (size_t c = 0; c < inner->n; c++) {
        size_t idx = inner->d[c];
        size_t k = idx / (Nx * Ny), j = (idx - k * Nx * Ny) / Nx, i = idx - k * Nx * Ny - j * Nx;
        if ((i + j + k) % 2 == 0) even.d[even.n++] = idx;
        else odd.d[odd.n++] = idx;
    }
    memcpy(out, in, size * sizeof(double)); /* out.copy_data(in), :41 / :114 */
    const double ax = Dc * dt * rdx * rdx, ay = Dc * dt * rdy * rdy, az = Dc * dt * rdz * rdz;
    const double rb_c = (1. + 2. * (ax + ay + az));
    const double beta = 1. / rb_c;
    int it = 0;
    double res = 1.;
    while (res > tol_res && it < max_iter) {
        memcpy(in, out, size * sizeof(double)); /* in.copy_data(out), :71 / :141 */
        const ivec *col[2] = {&even, &odd};
        for (int c = 0; c < 2; c++) {
            for (size_t n = 0; n < col[c]->n; n++) {
                size_t index = col[c]->d[n];
                if (!ev) {
                    cgs_stencil(index, out, b, ax, ay, az, dsign, beta, wgt, Nx, Ny);
                } else { /* turbulent step, :270-294 */
                    double aX = (Dc + ev[index]) * dt * rdx * rdx;
                    double aY = (Dc + ev[index]) * dt * rdy * rdy;
                    double aZ = (Dc + ev[index]) * dt * rdz * rdz;
                    double rb = (1. + 2. * (aX + aY + aZ));
                    double bb = 1. / rb;
                    cgs_stencil(index, out, b, aX, aY, aZ, dsign, bb, wgt, Nx, Ny);
                }
            }
        }
        or_apply_bc(D, out, ft, 0);
        double sum = 0;
        for (size_t l = 0; l < inner->n; ++l) {
            const size_t i = inner->d[l];
            double rb = rb_c;
            if (ev) {
                double aX = (Dc + ev[i]) * dt * rdx * rdx;
                double aY = (Dc + ev[i]) * dt * rdy * rdy;
                double aZ = (Dc + ev[i]) * dt * rdz * rdz;
                rb = (1. + 2. * (aX + aY + aZ));
            }
            double r = rb * (out[i] - in[i]);
            sum += r * r;
        }
        res = sqrt(sum);
        it++;
    }
    free(even.d);
    free(odd.d);
    if (res_out) *res_out = res;
    return it;
}

/* ---------------- divergence / projection, interfaces/IPressure.cpp ---------------- */
void or_divergence(const or_domain *D, double *out, const double *x, const double *y, const double *z) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1], Nz = D->N[0][2];
    const size_t nj = Nx, nk = Nx * Ny;
    const double rdx = 1. / D->h[0][0], rdy = 1. / D->h[0][1], rdz = 1. / D->h[0][2];
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        double vx = 0.5 * rdx * (x[i + 1] - x[i - 1]);
        double vy = 0.5 * rdy * (y[i + nj] - y[i - nj]);
        double vz = 0.5 * rdz * (z[i + nk] - z[i - nk]);
        out[i] = vx + vy + vz;
    }
    /* ghost shell set to zero, :69-104 */
    for (size_t k = 0; k < Nz; k++)
        for (size_t j = 0; j < Ny; j++)
            for (size_t i = 0; i < Nx; i++)
                if (i == 0 || j == 0 || k == 0 || i == Nx - 1 || j == Ny - 1 || k == Nz - 1)
                    out[i + Nx * j + Nx * Ny * k] = 0;
}

void or_projection(const or_domain *D, double *ou, double *ov, double *ow, const double *iu, const double *iv,
                   const double *iw, const double *ip) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const size_t nj = Nx, nk = Nx * Ny;
    const double rdx = 1. / D->h[0][0], rdy = 1. / D->h[0][1], rdz = 1. / D->h[0][2];
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        ou[i] = iu[i] - 0.5 * rdx * (ip[i + 1] - ip[i - 1]);
        ov[i] = iv[i] - 0.5 * rdy * (ip[i + nj] - ip[i - nj]);
        ow[i] = iw[i] - 0.5 * rdz * (ip[i + nk] - ip[i - nk]);
    }
    or_apply_bc(D, ou, OR_U, 0);
    or_apply_bc(D, ov, OR_V, 0);
    or_apply_bc(D, ow, OR_W, 0);
}

/* ---------------- ConstSmagorinsky, turbulence/ConstSmagorinsky.cpp:30-87 ---------------- */
void or_smagorinsky(const or_domain *D, double *ev, const double *u, const double *v, const double *w, double cs,
                    int ev_ft) {
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const size_t ni = 1, nj = Nx, nk = Nx * Ny;
    const double dx = D->h[0][0], dy = D->h[0][1], dz = D->h[0][2];
    const double rdx = 1. / dx, rdy = 1. / dy, rdz = 1. / dz;
    const double delta_s = cbrt(dx * dy * dz);
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        double S11 = (u[i + ni] - u[i - ni]) * 0.5 * rdx;
        double S22 = (v[i + nj] - v[i - nj]) * 0.5 * rdy;
        double S33 = (w[i + nk] - w[i - nk]) * 0.5 * rdz;
        double S12 = 0.5 * ((u[i + nj] - u[i - nj]) * 0.5 * rdy + (v[i + ni] - v[i - ni]) * 0.5 * rdx);
        double S13 = 0.5 * ((u[i + nk] - u[i - nk]) * 0.5 * rdz + (w[i + ni] - w[i - ni]) * 0.5 * rdx);
        double S23 = 0.5 * ((v[i + nk] - v[i - nk]) * 0.5 * rdz + (w[i + nj] - w[i - nj]) * 0.5 * rdy);
        double S_bar = sqrt(2. * (S11 * S11 + S22 * S22 + S33 * S33 + 2. * (S12 * S12) + 2. * (S13 * S13) +
                                  2. * (S23 * S23)));
        ev[i] = cs * cs * delta_s * delta_s * S_bar;
    }
    or_apply_bc(D, ev, ev_ft, 0);
}

/* ---------------- dynamic Smagorinsky ---------------- */
/* DynamicSmagorinsky::explicit_filtering, turbulence/DynamicSmagorinsky.cpp:405-427:
 * trapezoidal 3x3x3 filter on the inner cells, summed x offset outermost,
 * then y, then z. */
static void dyn_filter(const or_domain *D, double *out, const double *in) {
    const double a[3] = {1. / 4., 1. / 2., 1. / 4.};
    const long Nx = (long)D->N[0][0], NxNy = (long)(D->N[0][0] * D->N[0][1]);
    const ivec *inner = &D->inner[0];
    for (size_t q = 0; q < inner->n; ++q) {
        const long i = (long)inner->d[q];
        double sum = 0;
        for (long l = 0; l < 3; l++)
            for (long m = 0; m < 3; m++)
                for (long n = 0; n < 3; n++)
                    sum = sum + a[l] * a[m] * a[n] * in[i - 1 + l + Nx * (m - 1) + NxNy * (n - 1)];
        out[i] = sum;
    }
}

/* DynamicSmagorinsky::calc_turbulent_viscosity, DynamicSmagorinsky.cpp:172-393,
 * operation by operation.  The reference's work fields are never written
 * outside the inner cells and uw_f is never written at all (its filter call
 * filters ww twice, :235-240); here they are zero, the value a fresh
 * allocation reads.  The coefficient average (:372-384) runs in place in
 * inner-list order, so cells later in the list see averaged earlier ones. */
void or_dyn_smagorinsky(const or_domain *D, double *ev, const double *u, const double *v, const double *w,
                        double nu) {
    const size_t n = D->size[0];
    const long Nx = (long)D->N[0][0], NxNy = (long)(D->N[0][0] * D->N[0][1]);
    const double dx = D->h[0][0], dy = D->h[0][1], dz = D->h[0][2];
    const double rdx = 1. / dx, rdy = 1. / dy, rdz = 1. / dz;
    const double alpha = 2.0;
    const double delta_s = cbrt(dx * dy * dz);
    enum { UF, VF, WF, UU, VV, WW, UV, UW, VW, UU_F, VV_F, WW_F, UV_F, UW_F, VW_F, S11, S22, S33, S12, S13, S23,
           P11, P22, P33, P12, P13, P23, S11F, S22F, S33F, S12F, S13F, S23F, P11F, P22F, P33F, P12F, P13F, P23F,
           SBAR, CS, NWORK };
    double *f[NWORK];
    for (int q = 0; q < NWORK; q++) f[q] = (double *)calloc(n, sizeof(double));
    const ivec *inner = &D->inner[0];
    dyn_filter(D, f[UF], u);
    dyn_filter(D, f[VF], v);
    dyn_filter(D, f[WF], w);
    /* :212-228: both the "velocity products" and the products of the filtered
     * velocities are formed from the filtered velocities */
    for (size_t q = 0; q < inner->n; ++q) {
        const size_t i = inner->d[q];
        f[UU][i] = f[UF][i] * f[UF][i];
        f[VV][i] = f[VF][i] * f[VF][i];
        f[WW][i] = f[WF][i] * f[WF][i];
        f[UV][i] = f[UF][i] * f[VF][i];
        f[UW][i] = f[UF][i] * f[WF][i];
        f[VW][i] = f[VF][i] * f[WF][i];
    }
    dyn_filter(D, f[UU_F], f[UU]);
    dyn_filter(D, f[VV_F], f[VV]);
    dyn_filter(D, f[WW_F], f[WW]);
    dyn_filter(D, f[UV_F], f[UV]);
    dyn_filter(D, f[VW_F], f[VW]);
    /* uw_f: not filtered (stays zero) */
    for (size_t q = 0; q < inner->n; ++q) {
        const long i = (long)inner->d[q];
        const double *uf = f[UF], *vf = f[VF], *wf = f[WF];
        f[S11][i] = (uf[i + 1] - uf[i - 1]) * 0.5 * rdx;
        f[S22][i] = (vf[i + Nx] - vf[i - Nx]) * 0.5 * rdy;
        f[S33][i] = (wf[i + NxNy] - wf[i - NxNy]) * 0.5 * rdz;
        f[S12][i] = 0.5 * ((uf[i + Nx] - uf[i - Nx]) * 0.5 * rdy + (vf[i + 1] - vf[i - 1]) * 0.5 * rdx);
        f[S13][i] = 0.5 * ((uf[i + NxNy] - uf[i - NxNy]) * 0.5 * rdz + (wf[i + 1] - wf[i - 1]) * 0.5 * rdx);
        f[S23][i] = 0.5 * ((vf[i + NxNy] - vf[i - NxNy]) * 0.5 * rdz + (wf[i + Nx] - wf[i - Nx]) * 0.5 * rdy);
        f[SBAR][i] = sqrt(2. * (f[S11][i] * f[S11][i] + f[S22][i] * f[S22][i] + f[S33][i] * f[S33][i] +
                                2. * (f[S12][i] * f[S12][i]) + 2. * (f[S13][i] * f[S13][i]) +
                                2. * (f[S23][i] * f[S23][i])));
        for (int c = 0; c < 6; c++) f[P11 + c][i] = f[SBAR][i] * f[S11 + c][i];
    }
    for (int c = 0; c < 6; c++) dyn_filter(D, f[S11F + c], f[S11 + c]);
    for (int c = 0; c < 6; c++) dyn_filter(D, f[P11F + c], f[P11 + c]);
    for (size_t q = 0; q < inner->n; ++q) {
        const size_t i = inner->d[q];
        /* Leonard stress, :249-254 (uf_uf == uu and so on) */
        const double L11 = f[UU_F][i] - f[UU][i], L22 = f[VV_F][i] - f[VV][i], L33 = f[WW_F][i] - f[WW][i];
        const double L12 = f[UV_F][i] - f[UV][i], L13 = f[UW_F][i] - f[UW][i], L23 = f[VW_F][i] - f[VW][i];
        const double sf = sqrt(2. * (f[S11F][i] * f[S11F][i] + f[S22F][i] * f[S22F][i] + f[S33F][i] * f[S33F][i] +
                                     2. * (f[S12F][i] * f[S12F][i]) + 2. * (f[S13F][i] * f[S13F][i]) +
                                     2. * (f[S23F][i] * f[S23F][i])));
        double M[6];
        for (int c = 0; c < 6; c++)
            M[c] = 2.0 * delta_s * delta_s * (f[P11F + c][i] - alpha * alpha * sf * f[S11F + c][i]);
        /* :340-347, with the reference's products where sums belong */
        const double num = L11 * M[0] + L22 * M[1] * L33 * M[2] + 2.0 * L12 * M[3] + 2.0 * L13 * M[4] +
                           2.0 * L23 * M[5];
        const double den = M[0] * M[0] + M[1] * M[1] * M[2] * M[2] + 2.0 * M[3] * M[3] + 2.0 * M[4] * M[4] +
                           2.0 * M[5] * M[5];
        f[CS][i] = num / den;
    }
    for (size_t q = 0; q < inner->n; ++q) {
        const long i = (long)inner->d[q];
        double sum = 0;
        for (long l = 0; l < 3; l++)
            for (long m = 0; m < 3; m++)
                for (long nn = 0; nn < 3; nn++) sum = sum + f[CS][i - 1 + l + Nx * (m - 1) + NxNy * (nn - 1)];
        f[CS][i] = sum / 27.0;
        ev[i] = 2.0 * f[CS][i] * delta_s * delta_s * f[SBAR][i];
    }
    for (size_t q = 0; q < inner->n; ++q) {
        const size_t i = inner->d[q];
        if ((ev[i] + nu) < 0) ev[i] = 0;
    }
    for (int q = 0; q < NWORK; q++) free(f[q]);
}

/* ---------------- sources ---------------- */
/* ExplicitEulerSource::add_source, source/ExplicitEulerSource.cpp:84-104 */
void or_add_source(const or_domain *D, double *out, const double *s, double dt, int ft) {
    const ivec *inner = &D->inner[0];
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        out[i] += dt * s[i];
    }
    or_apply_bc(D, out, ft, 0);
}

/* ISource::buoyancy_force, interfaces/ISource.cpp:26-52 (domain list) */
void or_buoyancy(const or_domain *D, double *out, const double *T, const double *Ta, double beta, double g) {
    const ivec *dom = &D->domain[0];
    for (size_t l = 0; l < dom->n; ++l) {
        const size_t idx = dom->d[l];
        out[idx] = -beta * (T[idx] - Ta[idx]) * g;
    }
}

/* Obstacle::is_obstacle_cell with real coordinates, Obstacle.h:36-45 */
static int is_obst_cell_r(const obst_t *o, const double c[3]) {
    return (double)o->start[0] <= c[0] && c[0] <= (double)o->end[0] && (double)o->start[1] <= c[1] &&
           c[1] <= (double)o->end[1] && (double)o->start[2] <= c[2] && c[2] <= (double)o->end[2];
}
static int is_obst_cell_s(const obst_t *o, const size_t c[3]) {
    return o->start[0] <= c[0] && c[0] <= o->end[0] && o->start[1] <= c[1] && c[1] <= o->end[1] &&
           o->start[2] <= c[2] && c[2] <= o->end[2];
}

/* Obstacle::intersection, Obstacle.cpp:909-978 */
static int obst_intersection(const obst_t *o, const size_t src[3], const size_t cell[3]) {
    const double eps = 10E-6;
    if (is_obst_cell_s(o, cell)) return 1;
    double sl[3], el[3], dir[3];
    for (int a = 0; a < 3; a++) {
        sl[a] = (double)src[a] + 0.5;
        el[a] = (double)cell[a] + 0.5;
    }
    for (int a = 0; a < 3; a++) dir[a] = el[a] - sl[a];
    for (int axis = 0; axis < 3; axis++) {
        double sa[3], ea[3];
        for (int a = 0; a < 3; a++) {
            sa[a] = (double)o->start[a] + 0.5;
            ea[a] = (double)o->end[a] + 0.5;
        }
        for (int pp = 0; pp < 2; pp++) {
            int patch = 2 * axis + pp;
            if (patch % 2 == 1) {
                double t[3];
                memcpy(t, sa, sizeof(t));
                memcpy(sa, ea, sizeof(t));
                memcpy(ea, t, sizeof(t));
            }
            double nv[3] = {0, 0, 0};
            nv[axis] = 1;
            double res_dot = nv[0] * dir[0] + nv[1] * dir[1] + nv[2] * dir[2];
            if (fabs(res_dot) > eps) {
                double lambda = (sa[axis] - sl[axis]) / dir[axis];
                if (lambda <= 1 && lambda >= 0) {
                    double in[3];
                    for (int a = 0; a < 3; a++) in[a] = dir[a];
                    for (int a = 0; a < 3; a++) in[a] *= lambda;
                    for (int a = 0; a < 3; a++) in[a] += sl[a];
                    for (int a = 0; a < 3; a++) in[a] -= 0.5;
                    if (is_obst_cell_r(o, in)) return 1;
                }
            }
        }
    }
    return 0;
}

/* GaussFunction::create_spatial_values, source/GaussFunction.cpp:52-109 */
void or_gauss_spatial(const or_domain *D, double *out, const double *par) {
    const double HRR = par[0], cp = par[1];
    const double pos[3] = {par[2], par[3], par[4]};
    const double dim[3] = {par[5], par[6], par[7]};
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];
    const double sp[3] = {D->h[0][0], D->h[0][1], D->h[0][2]};
    double rs[3];
    for (int a = 0; a < 3; a++) {
        rs[a] = dim[a];
        rs[a] *= dim[a];
        rs[a] *= 2;
        rs[a] = 1. / rs[a];
    }
    size_t start[3];
    for (int a = 0; a < 3; a++) start[a] = (size_t)(round((-D->x0[a] + pos[a]) / sp[a]) + 1);
    memset(out, 0, D->size[0] * sizeof(double)); /* m_field_spatial_values(RHO, 0) */
    const ivec *inner = &D->inner[0];
    double V = 0.;
    for (int pass = 0; pass < 2; pass++) {
        double HRRrV = 0, rcp = 0;
        if (pass == 1) {
            HRRrV = HRR / V;
            rcp = 1. / cp;
        }
        for (size_t l = 0; l < inner->n; ++l) {
            const size_t idx = inner->d[l];
            size_t c[3];
            c[2] = idx / (Nx * Ny);
            c[1] = (idx - c[2] * Nx * Ny) / Nx;
            c[0] = idx - c[2] * Nx * Ny - c[1] * Nx;
            int blocked = 0;
            for (int id = 0; id < D->n_obst && !blocked; id++) blocked = obst_intersection(&D->obst[0][id], start, c);
            if (blocked) continue;
            double mid[3], d[3];
            for (int a = 0; a < 3; a++) mid[a] = D->x0[a] + ((double)c[a] - 0.5) * sp[a];
            for (int a = 0; a < 3; a++) d[a] = mid[a] - pos[a];
            for (int a = 0; a < 3; a++) d[a] *= d[a];
            for (int a = 0; a < 3; a++) d[a] *= rs[a];
            double expr = exp(-(d[0] + d[1] + d[2]));
            if (pass == 0) V += expr * sp[0] * sp[1] * sp[2];
            else out[idx] = HRRrV * rcp * expr;
        }
    }
}

/* ||b - lap_h x||_2 over the inner list of `level`, VCycleMG.cpp:186-195 */
double or_residual_norm(const or_domain *D, const double *x, const double *b, int level) {
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    const size_t nj = Nx, nk = Nx * Ny;
    const double dx = D->h[level][0], dy = D->h[level][1], dz = D->h[level][2];
    const double rdx2 = 1. / (dx * dx), rdy2 = 1. / (dy * dy), rdz2 = 1. / (dz * dz);
    const ivec *inner = &D->inner[level];
    double sum = 0.;
    for (size_t l = 0; l < inner->n; ++l) {
        const size_t i = inner->d[l];
        double r = b[i] - (rdx2 * (x[i - 1] - 2 * x[i] + x[i + 1]) + rdy2 * (x[i - nj] - 2 * x[i] + x[i + nj]) +
                           rdz2 * (x[i - nk] - 2 * x[i] + x[i + nk]));
        sum += r * r;
    }
    return sqrt(sum);
}
