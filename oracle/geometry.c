/* Oracle geometry: grid sizes per multigrid level, index lists and obstacle
 * face lists.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates domain/DomainData.{h,cpp}, domain/Domain.cpp, domain/Obstacle.cpp,
 * domain/Multigrid.cpp and utility/Algorithm.cpp of the reference.
 * Only the configuration the configs use is supported: computational domain
 * equal to the physical domain (enable_computational_domain="No"), no
 * surfaces (their list builder never terminates, Multigrid.cpp:292).
 */
#include "oracle_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- growable index vector ---------------- */
static void vpush(ivec *v, uint64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->d = (uint64_t *)realloc(v->d, v->cap * sizeof(uint64_t));
    }
    v->d[v->n++] = x;
}
static void vfree(ivec *v) { free(v->d); v->d = NULL; v->n = v->cap = 0; }

/* Algorithm::merge_sort, utility/Algorithm.cpp:22-56 (keeps duplicates) */
static ivec merge_keep(const uint64_t *a, size_t na, const uint64_t *b, size_t nb) {
    ivec r = {0};
    size_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) vpush(&r, a[i++]);
        else vpush(&r, b[j++]);
    }
    while (i < na) vpush(&r, a[i++]);
    while (j < nb) vpush(&r, b[j++]);
    return r;
}

/* Algorithm::merge_sort_with_duplicates, utility/Algorithm.cpp:67-113.
 * Both inputs sorted and non-empty; equal heads are emitted once. */
static ivec merge_unique(const uint64_t *a, size_t na, const uint64_t *b, size_t nb) {
    ivec r = {0};
    size_t i = 0, j = 0;
    if (na == 0 || nb == 0) { /* not reached by the reference's call sites */
        for (; i < na; i++) vpush(&r, a[i]);
        for (; j < nb; j++) vpush(&r, b[j]);
        return r;
    }
    while (i < na && j < nb) {
        if (a[i] < b[j]) vpush(&r, a[i++]);
        else if (a[i] > b[j]) vpush(&r, b[j++]);
        else { vpush(&r, a[i]); i++; j++; }
    }
    if (i < na) {
        if (a[i] == r.d[r.n - 1]) i++;
        for (; i < na; i++) vpush(&r, a[i]);
    } else if (j < nb) {
        if (b[j] == r.d[r.n - 1]) j++;
        for (; j < nb; j++) vpush(&r, b[j]);
    }
    return r;
}

static size_t IXs(size_t i, size_t j, size_t k, size_t Nx, size_t Ny) { return i + Nx * j + Nx * Ny * k; }

/* ---------------- DomainData (DomainData.cpp:12-56, DomainData.h:144-171) ---------------- */
static double spacing(const or_domain *D, int axis, int level) {
    return D->len[axis] / (double)D->n[level][axis];
}
static size_t start_index(const or_domain *D, int axis, int level) {
    /* round((start_CD - start_PD) / h) + 1 with CD == PD */
    return (size_t)round((D->x0[axis] - D->x0[axis]) / spacing(D, axis, level)) + 1;
}
static size_t end_index(const or_domain *D, int axis, int level) {
    return (size_t)round((D->x1[axis] - D->x0[axis]) / spacing(D, axis, level));
}

/* ---------------- Obstacle (Obstacle.cpp:13-191, 449-478) ---------------- */
static void obstacle_init(or_domain *D, obst_t *o, int level) {
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    size_t sx = o->end[0] - o->start[0] + 1, sy = o->end[1] - o->start[1] + 1, sz = o->end[2] - o->start[2] + 1;
    size_t nface[6];
    nface[OR_FRONT] = sy * sx;
    nface[OR_BACK] = sy * sx;
    nface[OR_BOTTOM] = sz * sx;
    nface[OR_TOP] = sz * sx;
    nface[OR_LEFT] = sz * sy;
    nface[OR_RIGHT] = sz * sy;
    /* remove_cells_at_boundary, Obstacle.cpp:458-478 */
    if (o->start[2] <= start_index(D, 2, level)) nface[OR_FRONT] = 0;
    if (o->end[2] >= end_index(D, 2, level)) nface[OR_BACK] = 0;
    if (o->start[1] <= start_index(D, 1, level)) nface[OR_BOTTOM] = 0;
    if (o->end[1] >= end_index(D, 1, level)) nface[OR_TOP] = 0;
    if (o->start[0] <= start_index(D, 0, level)) nface[OR_LEFT] = 0;
    if (o->end[0] >= end_index(D, 0, level)) nface[OR_RIGHT] = 0;

    /* create_obstacle, Obstacle.cpp:84-191 */
    memset(&o->cells, 0, sizeof(ivec));
    for (size_t k = o->start[2]; k <= o->end[2]; ++k)
        for (size_t j = o->start[1]; j <= o->end[1]; ++j)
            for (size_t i = o->start[0]; i <= o->end[0]; ++i) vpush(&o->cells, IXs(i, j, k, Nx, Ny));
    const uint64_t *c = o->cells.d;
    for (int p = 0; p < 6; p++) memset(&o->face[p], 0, sizeof(ivec));
    if (nface[OR_FRONT] > 0)
        for (size_t j = 0; j < sy; ++j)
            for (size_t i = 0; i < sx; ++i) vpush(&o->face[OR_FRONT], c[IXs(i, j, 0, sx, sy)]);
    if (nface[OR_BACK] > 0)
        for (size_t j = 0; j < sy; ++j)
            for (size_t i = 0; i < sx; ++i) vpush(&o->face[OR_BACK], c[IXs(i, j, sz - 1, sx, sy)]);
    if (nface[OR_BOTTOM] > 0)
        for (size_t k = 0; k < sz; ++k)
            for (size_t i = 0; i < sx; ++i) vpush(&o->face[OR_BOTTOM], c[IXs(i, 0, k, sx, sy)]);
    if (nface[OR_TOP] > 0)
        for (size_t k = 0; k < sz; ++k)
            for (size_t i = 0; i < sx; ++i) vpush(&o->face[OR_TOP], c[IXs(i, sy - 1, k, sx, sy)]);
    if (nface[OR_LEFT] > 0)
        for (size_t k = 0; k < sz; ++k)
            for (size_t j = 0; j < sy; ++j) vpush(&o->face[OR_LEFT], c[IXs(0, j, k, sx, sy)]);
    if (nface[OR_RIGHT] > 0)
        for (size_t k = 0; k < sz; ++k)
            for (size_t j = 0; j < sy; ++j) vpush(&o->face[OR_RIGHT], c[IXs(sx - 1, j, k, sx, sy)]);
}

/* Obstacle::calculate_area_index, Obstacle.cpp:535-559 */
static void area_index(const obst_t *o1, const obst_t *o2, size_t *c1, size_t *c2, int axis, int start) {
    if (start) {
        *c1 = o1->start[axis];
        *c2 = o2->start[axis];
        if (o1->start[axis] > o2->start[axis]) *c2 = o1->start[axis] + 1;
        else if (o1->start[axis] < o2->start[axis]) *c1 = o2->start[axis] + 1;
    } else {
        *c1 = o1->end[axis];
        *c2 = o2->end[axis];
        if (o1->end[axis] < o2->end[axis]) *c2 = o1->end[axis] - 1;
        else if (o1->end[axis] > o2->end[axis]) *c1 = o2->end[axis] - 1;
    }
}

static int has_overlap1(size_t a1, size_t a2, size_t b1, size_t b2) { return a1 <= b2 && a2 >= b1; }

/* One side of the face-list walk of Obstacle::circular_constraints
 * (Obstacle.cpp:655-771): the walk removes the cells enumerated from the
 * removal area's start corner, a0 fastest, advancing only on a match. */
typedef struct {
    const uint64_t *old;
    size_t n, pos;
    ivec out;
    size_t ax0, ax1, rs[3], re[3];
    size_t cur0, cur1, removing;
    int ended;
    size_t Nx, Ny, axis;
} walk_t;

static size_t coord_index(const size_t c[3], size_t Nx, size_t Ny) { return IXs(c[0], c[1], c[2], Nx, Ny); }

static void walk_advance_removal(walk_t *w, int check_end) {
    w->cur0++;
    if (w->cur0 > w->re[w->ax0]) {
        w->cur0 = w->rs[w->ax0];
        w->cur1++;
        if (check_end && w->cur1 > w->re[w->ax1]) w->ended = 1;
    }
    size_t t[3];
    t[w->ax0] = w->cur0;
    t[w->ax1] = w->cur1;
    t[w->axis] = w->rs[w->axis];
    w->removing = coord_index(t, w->Nx, w->Ny);
}

static void circular_constraints(or_domain *D, obst_t *o1, obst_t *o2, int axis) {
    if (o1->start[axis] - 1 != o2->end[axis]) return;
    int a0, a1;
    if (axis == 0) { a0 = 1; a1 = 2; }
    else if (axis == 1) { a0 = 0; a1 = 2; }
    else { a0 = 0; a1 = 1; }
    if (!(has_overlap1(o1->start[a0], o1->end[a0], o2->start[a0], o2->end[a0]) &&
          has_overlap1(o1->start[a1], o1->end[a1], o2->start[a1], o2->end[a1])))
        return;
    const int p1 = 2 * axis;     /* Mapping::to_patch(axis, true): LEFT/BOTTOM/FRONT */
    const int p2 = 2 * axis + 1; /* RIGHT/TOP/BACK */
    const size_t Nx = D->N[0][0], Ny = D->N[0][1];

    walk_t w[2];
    memset(w, 0, sizeof(w));
    obst_t *ob[2] = {o1, o2};
    int pt[2] = {p1, p2};
    w[0].rs[axis] = w[0].re[axis] = o1->start[axis];
    w[1].rs[axis] = w[1].re[axis] = o2->end[axis];
    size_t c1, c2;
    area_index(o1, o2, &c1, &c2, a0, 1); w[0].rs[a0] = c1; w[1].rs[a0] = c2;
    area_index(o1, o2, &c1, &c2, a0, 0); w[0].re[a0] = c1; w[1].re[a0] = c2;
    area_index(o1, o2, &c1, &c2, a1, 1); w[0].rs[a1] = c1; w[1].rs[a1] = c2;
    area_index(o1, o2, &c1, &c2, a1, 0); w[0].re[a1] = c1; w[1].re[a1] = c2;

    for (int s = 0; s < 2; s++) {
        walk_t *x = &w[s];
        x->old = ob[s]->face[pt[s]].d;
        x->n = ob[s]->face[pt[s]].n;
        x->ax0 = (size_t)a0; x->ax1 = (size_t)a1; x->axis = (size_t)axis;
        x->Nx = Nx; x->Ny = Ny;
        x->removing = coord_index(x->rs, Nx, Ny);
        x->cur0 = x->rs[a0];
        x->cur1 = x->rs[a1];
        /* copy everything below the smallest removal index (Obstacle.cpp:683-703) */
        while (x->pos < x->n && x->old[x->pos] < x->removing) vpush(&x->out, x->old[x->pos++]);
    }
    /* joint walk, Obstacle.cpp:713-756 */
    for (; w[0].pos < w[0].n && w[1].pos < w[1].n && !w[0].ended && !w[1].ended; w[0].pos++, w[1].pos++) {
        for (int s = 0; s < 2; s++) {
            uint64_t cur = w[s].old[w[s].pos];
            if (cur != w[s].removing) vpush(&w[s].out, cur);
            else walk_advance_removal(&w[s], 1);
        }
    }
    for (int s = 0; s < 2; s++) {
        walk_t *x = &w[s];
        /* Obstacle.cpp:758-795: continue alone while inside the removal rows */
        if (!x->ended) {
            for (; x->pos < x->n && x->cur1 <= x->re[x->ax1]; x->pos++) {
                uint64_t cur = x->old[x->pos];
                if (cur != x->removing) vpush(&x->out, cur);
                else walk_advance_removal(x, 0);
            }
        }
        for (; x->pos < x->n; x->pos++) vpush(&x->out, x->old[x->pos]);
        vfree(&ob[s]->face[pt[s]]);
        ob[s]->face[pt[s]] = x->out;
    }
}

/* Obstacle::remove_circular_constraints, Obstacle.cpp:562-590 */
static void remove_circular_constraints(or_domain *D, obst_t *o1, obst_t *o2) {
    for (int axis = 0; axis < 3; axis++) {
        if (o1->end[axis] + 1 == o2->start[axis]) circular_constraints(D, o2, o1, axis);
        else circular_constraints(D, o1, o2, axis);
    }
}

/* ---------------- Domain lists (Domain.cpp:36-55, 94-172, 180-210, 241-252) ---------------- */
static void build_domain_lists(or_domain *D, int level) {
    const size_t Nx = D->N[level][0], Ny = D->N[level][1];
    const ivec *ob = &D->obst_cells[level];
    size_t oc = 0;
    ivec *inner = &D->inner[level];
    memset(inner, 0, sizeof(ivec));
    for (size_t k = start_index(D, 2, level); k <= end_index(D, 2, level); ++k)
        for (size_t j = start_index(D, 1, level); j <= end_index(D, 1, level); ++j)
            for (size_t i = start_index(D, 0, level); i <= end_index(D, 0, level); ++i) {
                size_t idx = IXs(i, j, k, Nx, Ny);
                if (oc < ob->n && ob->d[oc] == idx) oc++;
                else vpush(inner, idx);
            }
    /* boundary_cells, Domain.cpp:94-172 */
    for (int axis = 0; axis < 3; axis++) {
        int ax1, ax2;
        if (axis == 0) { ax1 = 1; ax2 = 2; }
        else if (axis == 1) { ax1 = 0; ax2 = 2; }
        else { ax1 = 0; ax2 = 1; }
        const int ps = 2 * axis, pe = 2 * axis + 1;
        memset(&D->bnd[level][ps], 0, sizeof(ivec));
        memset(&D->bnd[level][pe], 0, sizeof(ivec));
        size_t ts[3], te[3];
        size_t z1 = start_index(D, ax2, level) - 1, y1 = start_index(D, ax1, level) - 1;
        size_t z2 = end_index(D, ax2, level) + 1, y2 = end_index(D, ax1, level) + 1;
        ts[axis] = start_index(D, axis, level) - 1;
        te[axis] = end_index(D, axis, level) + 1;
        for (size_t k = z1; k <= z2; k++) {
            ts[ax2] = k;
            te[ax2] = k;
            for (size_t j = y1; j <= y2; j++) {
                ts[ax1] = j;
                vpush(&D->bnd[level][ps], coord_index(ts, Nx, Ny));
                te[ax1] = j;
                vpush(&D->bnd[level][pe], coord_index(te, Nx, Ny));
            }
        }
    }
    /* joined_list, Domain.cpp:241-252 */
    ivec b = {0};
    for (size_t i = 0; i < D->bnd[level][0].n; i++) vpush(&b, D->bnd[level][0].d[i]);
    for (int p = 1; p < 6; p++) {
        ivec m = merge_unique(b.d, b.n, D->bnd[level][p].d, D->bnd[level][p].n);
        vfree(&b);
        b = m;
    }
    D->domain[level] = merge_keep(inner->d, inner->n, b.d, b.n);
    vfree(&b);
}

or_domain *or_domain_create(const or_desc *d, char *err, size_t errlen) {
    or_domain *D = (or_domain *)calloc(1, sizeof(or_domain));
    D->desc = *d;
    D->L = d->levels;
    if (D->L < 0 || D->L >= OR_MAX_LEVELS || d->n_obst < 0 || d->n_obst > OR_MAX_OBST) {
        snprintf(err, errlen, "levels/obstacles out of range");
        free(D);
        return NULL;
    }
    D->x0[0] = d->X1; D->x1[0] = d->X2;
    D->x0[1] = d->Y1; D->x1[1] = d->Y2;
    D->x0[2] = d->Z1; D->x1[2] = d->Z2;
    for (int a = 0; a < 3; a++) D->len[a] = fabs(D->x1[a] - D->x0[a]);
    D->n[0][0] = (size_t)d->nx; D->n[0][1] = (size_t)d->ny; D->n[0][2] = (size_t)d->nz;
    for (int l = 0; l <= D->L; l++) {
        for (int a = 0; a < 3; a++) {
            if (l > 0) D->n[l][a] = (D->n[l - 1][a] == 1) ? 1 : (size_t)round((double)(D->n[l - 1][a] / 2));
            D->N[l][a] = (size_t)(round(D->len[a] / spacing(D, a, l)) + 2);
        }
        D->size[l] = D->N[l][0] * D->N[l][1] * D->N[l][2];
        for (int a = 0; a < 3; a++) D->h[l][a] = spacing(D, a, l);
    }
    D->n_obst = d->n_obst;
    /* level 0 obstacles from physical coordinates, Obstacle.cpp:13-30 */
    for (int id = 0; id < d->n_obst; id++) {
        obst_t *o = &D->obst[0][id];
        for (int a = 0; a < 3; a++) {
            double h = spacing(D, a, 0);
            int s = (int)round((-D->x0[a] + d->obst[id][2 * a]) / h);
            int e = (int)round((-D->x0[a] + d->obst[id][2 * a + 1]) / h);
            if (s + 1 < 1 || e > (int)D->n[0][a] || s + 1 > e) {
                snprintf(err, errlen, "obstacle %d does not lie inside the domain", id);
                or_domain_free(D);
                return NULL;
            }
            o->start[a] = (size_t)(s + 1);
            o->end[a] = (size_t)e;
        }
        obstacle_init(D, o, 0);
    }
    /* overlapping level-0 obstacles make the reference's inner list overflow */
    for (int i = 0; i < d->n_obst; i++)
        for (int j = i + 1; j < d->n_obst; j++) {
            int ov = 1;
            for (int a = 0; a < 3; a++)
                ov &= has_overlap1(D->obst[0][i].start[a], D->obst[0][i].end[a], D->obst[0][j].start[a],
                                   D->obst[0][j].end[a]);
            if (ov) {
                snprintf(err, errlen, "obstacles %d and %d overlap", i, j);
                or_domain_free(D);
                return NULL;
            }
        }
    /* DomainController::detect_neighbouring_obstacles, DomainController.cpp:150-165 */
    for (int i = 0; i < d->n_obst; i++)
        for (int j = i + 1; j < d->n_obst; j++) remove_circular_constraints(D, &D->obst[0][i], &D->obst[0][j]);
    /* Multigrid::create_multigrid_obstacle_lists / obstacle_dominant_restriction */
    for (int l = 0; l <= D->L; l++) {
        if (l > 0) {
            for (int id = 0; id < d->n_obst; id++) {
                obst_t *o = &D->obst[l][id];
                for (int a = 0; a < 3; a++) {
                    o->start[a] = (D->obst[l - 1][id].start[a] + 1) / 2;
                    o->end[a] = (D->obst[l - 1][id].end[a] + 1) / 2;
                }
                obstacle_init(D, o, l);
            }
        }
        memset(&D->obst_cells[l], 0, sizeof(ivec));
        if (d->n_obst > 0) {
            ivec m = {0};
            for (size_t i = 0; i < D->obst[l][0].cells.n; i++) vpush(&m, D->obst[l][0].cells.d[i]);
            for (int id = 1; id < d->n_obst; id++) {
                ivec t = (l == 0) ? merge_keep(m.d, m.n, D->obst[l][id].cells.d, D->obst[l][id].cells.n)
                                  : merge_unique(m.d, m.n, D->obst[l][id].cells.d, D->obst[l][id].cells.n);
                vfree(&m);
                m = t;
            }
            D->obst_cells[l] = m;
        }
        build_domain_lists(D, l);
    }
    return D;
}

void or_domain_free(or_domain *D) {
    if (!D) return;
    for (int l = 0; l <= D->L && l < OR_MAX_LEVELS; l++) {
        vfree(&D->inner[l]);
        vfree(&D->domain[l]);
        vfree(&D->obst_cells[l]);
        for (int p = 0; p < 6; p++) vfree(&D->bnd[l][p]);
        for (int id = 0; id < D->n_obst; id++) {
            vfree(&D->obst[l][id].cells);
            for (int p = 0; p < 6; p++) vfree(&D->obst[l][id].face[p]);
        }
    }
    free(D);
}

void or_level_info(const or_domain *D, int level, size_t *o7, double *h3) {
    o7[0] = D->N[level][0]; o7[1] = D->N[level][1]; o7[2] = D->N[level][2];
    o7[3] = D->n[level][0]; o7[4] = D->n[level][1]; o7[5] = D->n[level][2];
    o7[6] = D->size[level];
    for (int a = 0; a < 3; a++) h3[a] = D->h[level][a];
}

size_t or_list(const or_domain *D, int which, int level, int id, const uint64_t **out) {
    const ivec *v = NULL;
    if (which == 0) v = &D->inner[level];
    else if (which == 1) v = &D->domain[level];
    else if (which == 2) v = &D->obst_cells[level];
    else if (which >= 3 && which < 9) v = &D->bnd[level][which - 3];
    else if (which >= 10 && which < 16) v = &D->obst[level][id].face[which - 10];
    if (!v) { *out = NULL; return 0; }
    *out = v->d;
    return v->n;
}

void or_obstacle_box(const or_domain *D, int level, int id, size_t *s3, size_t *e3) {
    for (int a = 0; a < 3; a++) {
        s3[a] = D->obst[level][id].start[a];
        e3[a] = D->obst[level][id].end[a];
    }
}
