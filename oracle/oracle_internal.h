/* Internal structures of the CPU oracle.  TEST INFRASTRUCTURE ONLY. */
#ifndef ARTSS_ORACLE_INTERNAL_H
#define ARTSS_ORACLE_INTERNAL_H

#include "oracle.h"

typedef struct {
    uint64_t *d;
    size_t n, cap;
} ivec;

typedef struct {
    size_t start[3], end[3];
    ivec cells;    /* Obstacle::m_obstacle_list */
    ivec face[6];  /* Obstacle::m_boundary[patch] */
} obst_t;

struct or_domain {
    or_desc desc;
    int L;
    double x0[3], x1[3], len[3];
    size_t n[OR_MAX_LEVELS][3]; /* inner cells per axis */
    size_t N[OR_MAX_LEVELS][3]; /* cells incl. ghosts */
    size_t size[OR_MAX_LEVELS];
    double h[OR_MAX_LEVELS][3];
    ivec inner[OR_MAX_LEVELS];
    ivec domain[OR_MAX_LEVELS];
    ivec obst_cells[OR_MAX_LEVELS];
    ivec bnd[OR_MAX_LEVELS][6];
    int n_obst;
    obst_t obst[OR_MAX_LEVELS][OR_MAX_OBST];
};

#endif
