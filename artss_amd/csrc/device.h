// Device-side descriptors shared by the HIP kernels and the host code.
//
// Field layout is the reference's (utility/GlobalMacrosTypes.h:70): float64,
// x fastest, idx = i + Nx*j + Nx*Ny*k, one ghost layer, so a device array can
// be handed to / taken from a reference Field without reordering.
//
// The reference iterates 8-byte index lists (GPULists/SingleJoinedList).
// Here every level is a structured box and a cell is an inner cell when it
// lies in [1,n] on each axis and in none of the level's obstacle boxes, which
// is exactly the reference's inner list (Domain.cpp:180-210 subtracts the
// merged obstacle list, Multigrid.cpp:438-514 restricts obstacles as boxes).
// Kernels therefore read no index list at all and access memory coalesced
// along x.
#pragma once

#include <cstdint>

namespace artss {

constexpr int kMaxObstacles = 32;  // obstacle boxes per level carried in kernel arguments

struct Box {
    int lo[3];
    int hi[3];  // inclusive
};

struct LevelDesc {
    int N[3];       // cells including ghosts (Nx, Ny, Nz)
    int n[3];       // inner cells
    long long size; // Nx*Ny*Nz
    double h[3];    // spacing
    int n_obst;
    Box obst[kMaxObstacles];
};

// Domain ghost fill for one (field type, level): per patch an affine map
// g = sign * f[ref] + value.  The kernel evaluates, per ghost cell, the chain
// of patches that write it in the reference order LEFT, RIGHT, BOTTOM, TOP,
// FRONT, BACK (DomainBoundary.cpp:193-239), so edges and corners get exactly
// the value the sequential patch loop leaves there.
struct DomainBC {
    int active;        // field type has boundary data (BoundaryData::is_empty)
    double sign[6];
    double value[6];
    int target[6];     // coordinate along the patch axis of the referenced cell
};

// One obstacle face write: f[idx] = sign * f[idx + ref] + value, parameters
// from a per-(field type, level) table indexed by `param`.
struct FaceEntry {
    uint32_t idx;
    int32_t ref;
    int32_t param;  // obstacle * 6 + patch
    int32_t pad;
};

struct FaceParam {
    double sign;
    double value;
    int active;  // 0: no write for this field type (no data or periodic)
    int pad;
};

// A ghost cell whose patch chain starts on an obstacle face cell, with the
// face write (index into the face entries) that produces that cell.
struct DepGhost {
    uint32_t idx;
    int32_t face;
};

// Everything a kernel needs to produce the values of an idempotent boundary
// fill itself (faces of one wave, ghosts, ghosts depending on faces).
// Outcome of one residual-target loop of a captured step, written on the
// device: the iterations that ran, whether the loop ended as its eager form
// would (ok), and the final residual
struct LoopLog {
    double value;
    int count;
    int ok;
    double start;  // the residual at the loop's start, where it has one
};

struct BCView {
    const FaceEntry *faces;
    long long n_faces;
    const FaceParam *params;
    const DepGhost *deps;
    long long n_deps;
    const uint32_t *wbits;  // bit c set: cell c is written by the fill
    DomainBC bc;
    long long n_ghosts;     // 0 when the field type has no domain boundary data
    // ghost cells a sweep kernel's main grid does not write itself (edges,
    // corners, face ghosts whose chain starts off the inner fluid), without
    // the dependent ones
    const uint32_t *rest;
    long long n_rest;
    // every ghost cell is written by the fill: kernels that must write every
    // cell (first sweeps, the out-of-place prolongation, copy_bc) can then
    // run over the inner box only
    int shell_filled;
};

#if defined(__HIPCC__)
__host__ __device__ inline bool in_box(const Box &b, int i, int j, int k) {
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2];
}
// The first two boxes are tested without a data-dependent loop so that their
// bounds are fetched together, not one scalar load at a time.
__host__ __device__ inline bool in_obstacle(const LevelDesc &L, int i, int j, int k) {
    const int n = L.n_obst;
    bool r = (n > 0 && in_box(L.obst[0], i, j, k)) | (n > 1 && in_box(L.obst[1], i, j, k));
    for (int o = 2; o < n; o++) r = r || in_box(L.obst[o], i, j, k);
    return r;
}
#endif

}  // namespace artss
