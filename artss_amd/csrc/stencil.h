// Device-side cell operations shared by every kernel that applies them.
//
// The multi-workgroup kernels (kernels.hip) and the single-workgroup
// coarse-level multigrid kernel (mg_tail.hip) evaluate the same expressions
// through these functions, so both produce bit-identical values; each keeps
// the reference's operand order (files built with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace artss {
namespace dev {

__device__ inline long long lin(const LevelDesc &L, long long i, long long j, long long k) {
    return i + static_cast<long long>(L.N[0]) * (j + static_cast<long long>(L.N[1]) * k);
}

// VCycleMG.cpp:346-354 (Residuum) and :186-193 (residual norm): discrete
// Laplacian with 1/h^2 formed in place, as the reference does (in T).
template <typename T>
__device__ inline T laplace_t(const LevelDesc &L, const T *__restrict__ x, long long i) {
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const T rdx2 = static_cast<T>(1. / (L.h[0] * L.h[0])), rdy2 = static_cast<T>(1. / (L.h[1] * L.h[1])),
            rdz2 = static_cast<T>(1. / (L.h[2] * L.h[2]));
    return (rdx2 * (x[i - 1] - 2 * x[i] + x[i + 1]) + rdy2 * (x[i - nj] - 2 * x[i] + x[i + nj]) +
            rdz2 * (x[i - nk] - 2 * x[i] + x[i + nk]));
}
__device__ inline double laplace(const LevelDesc &L, const double *__restrict__ x, long long i) {
    return laplace_t<double>(L, x, i);
}

// JacobiDiffuse.cpp:292-300: one weighted Jacobi update of cell idx, in the
// arithmetic of the iterate's type T (double: the reference's bits; float:
// the mixed-precision V-cycle); the right-hand side may be stored wider.
template <typename T, typename TB>
__device__ inline T jacobi_t(const LevelDesc &L, T ax, T ay, T az, T beta, T dsign, T w, const T *__restrict__ in,
                             const TB *__restrict__ b, long long idx) {
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const T out_h = beta * (dsign * static_cast<T>(b[idx]) + ax * (in[idx + 1] + in[idx - 1]) +
                            ay * (in[idx + nj] + in[idx - nj]) + az * (in[idx + nk] + in[idx - nk]));
    return (1 - w) * in[idx] + w * out_h;
}
__device__ inline double jacobi_mg(const LevelDesc &L, double ax, double ay, double az, double beta, double dsign,
                                   double w, const double *__restrict__ in, const double *__restrict__ b,
                                   long long idx) {
    return jacobi_t<double, double>(L, ax, ay, az, beta, dsign, w, in, b, idx);
}

// ColoredGaussSeidelDiffuse.cpp:316-341, constant coefficients, in place
__device__ inline double cgs_mg(const LevelDesc &L, double ax, double ay, double az, double beta, double dsign,
                                double w, const double *out, const double *__restrict__ b, long long index) {
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const double ox = out[index + 1], ox2 = out[index - 1];
    const double oy = out[index + nj], oy2 = out[index - nj];
    const double oz = out[index + nk], oz2 = out[index - nk];
    const double db = b[index], r_out = out[index];
    const double out_h = beta * (dsign * db + ax * (ox + ox2) + ay * (oy + oy2) + az * (oz + oz2));
    return (1 - w) * r_out + w * out_h;
}

// VCycleMG.cpp:404-421: mean of the eight fine cells of coarse cell (i,j,k)
template <typename T>
__device__ inline T restrict_t(const LevelDesc &Lf, const T *__restrict__ in, long long i, long long j, long long k) {
    const long long i2 = 2 * i, j2 = 2 * j, k2 = 2 * k;
    return static_cast<T>(0.125) *
           (in[lin(Lf, i2 - 1, j2 - 1, k2 - 1)] + in[lin(Lf, i2, j2 - 1, k2 - 1)] + in[lin(Lf, i2 - 1, j2, k2 - 1)] +
            in[lin(Lf, i2, j2, k2 - 1)] + in[lin(Lf, i2 - 1, j2 - 1, k2)] + in[lin(Lf, i2, j2 - 1, k2)] +
            in[lin(Lf, i2 - 1, j2, k2)] + in[lin(Lf, i2, j2, k2)]);
}
__device__ inline double restrict_cell(const LevelDesc &Lf, const double *__restrict__ in, long long i, long long j,
                                       long long k) {
    return restrict_t<double>(Lf, in, i, j, k);
}

// VCycleMG.cpp:466-546: value of fine cell number q (0..7) of coarse cell idx,
// weights (27, 9, 9, 9, 3, 3, 3, 1) / 64 toward the neighbours on that
// fine cell's side.  q's bits select the -1 side along x (4), y (2), z (1);
// q = 0..7 is the reference's order of the eight writes.
template <typename T>
__device__ inline T prolong_t(const LevelDesc &Lc, const T *__restrict__ in, long long idx, int q) {
    const long long ni = 1, nj = Lc.N[0], nk = static_cast<long long>(Lc.N[0]) * Lc.N[1];
    const int di = (q & 4) ? -1 : 1, dj = (q & 2) ? -1 : 1, dk = (q & 1) ? -1 : 1;
    const long long oi = di * ni, oj = dj * nj, ok = dk * nk;
    const T c0 = in[idx];
    return static_cast<T>(0.015625) * (27 * c0 + 9 * in[idx + oi] + 9 * in[idx + oj] + 9 * in[idx + ok] +
                                       3 * in[idx + oi + oj] + 3 * in[idx + oi + ok] + 3 * in[idx + oj + ok] +
                                       in[idx + oi + oj + ok]);
}
__device__ inline double prolong_value(const LevelDesc &Lc, const double *__restrict__ in, long long idx, int q) {
    return prolong_t<double>(Lc, in, idx, q);
}
__device__ inline long long prolong_target(const LevelDesc &Lf, long long i, long long j, long long k, int q) {
    return lin(Lf, 2 * i - (q & 4 ? 1 : 0), 2 * j - (q & 2 ? 1 : 0), 2 * k - (q & 1 ? 1 : 0));
}

// ObstacleBoundary.cpp:42-48: f[idx] = sign * f[idx +- ref] + value
__device__ inline void face_write(const FaceEntry &x, const FaceParam *__restrict__ p, double *f) {
    const FaceParam q = p[x.param];
    if (!q.active) return;
    f[x.idx] = q.sign * f[static_cast<long long>(x.idx) + x.ref] + q.value;
}

// number of ghost cells of a level (the six faces of the shell, edges and
// corners counted once)
__device__ __host__ inline long long ghost_count(const LevelDesc &L) {
    const long long Nx = L.N[0], Ny = L.N[1], Nz = L.N[2];
    return 2 * Ny * Nz + 2 * (Nx - 2) * Nz + 2 * (Nx - 2) * (Ny - 2);
}

// Ghost cell number t (0 <= t < ghost_count) of a level: its coordinates
// (the x faces first, then y without x edges, then z without x and y edges)
// (32-bit arithmetic: a level's ghost count stays below 2^31)
// (the coordinates are formed in locals and stored once: stores to the
// reference parameters from different branches make the compiler keep them
// in per-lane scratch memory)
__device__ inline bool ghost_coords(const LevelDesc &L, long long t64, long long &i, long long &j, long long &k) {
    const int Nx = L.N[0], Ny = L.N[1], Nz = L.N[2];
    const int fx = Ny * Nz, fy = (Nx - 2) * Nz, fz = (Nx - 2) * (Ny - 2);
    int t = static_cast<int>(t64);
    int a, b, c;  // i, j, k
    if (t < 2 * fx) {
        const int side = t >= fx, r = side ? t - fx : t;
        a = side ? Nx - 1 : 0;
        b = r % Ny;
        c = r / Ny;
    } else if ((t -= 2 * fx) < 2 * fy) {
        const int side = t >= fy, r = side ? t - fy : t;
        b = side ? Ny - 1 : 0;
        a = 1 + r % (Nx - 2);
        c = r / (Nx - 2);
    } else if ((t -= 2 * fy) < 2 * fz) {
        const int side = t >= fz, r = side ? t - fz : t;
        c = side ? Nz - 1 : 0;
        a = 1 + r % (Nx - 2);
        b = 1 + r / (Nx - 2);
    } else {
        return false;
    }
    i = a;
    j = b;
    k = c;
    return true;
}

// the cell a ghost's patch chain reads, and the chain applied to its value
__device__ inline long long ghost_source(const LevelDesc &L, const DomainBC &bc, long long i, long long j,
                                         long long k) {
    const long long Nx = L.N[0], Ny = L.N[1], Nz = L.N[2];
    const long long ii = (i == 0 || i == Nx - 1) ? bc.target[i == 0 ? 0 : 1] : i;
    const long long jj = (j == 0 || j == Ny - 1) ? bc.target[j == 0 ? 2 : 3] : j;
    const long long kk = (k == 0 || k == Nz - 1) ? bc.target[k == 0 ? 4 : 5] : k;
    return ii + Nx * (jj + Ny * kk);
}
template <typename T = double>
__device__ inline T ghost_chain(const LevelDesc &L, const DomainBC &bc, long long i, long long j, long long k, T v) {
    const long long Nx = L.N[0], Ny = L.N[1], Nz = L.N[2];
    const int px = i == 0 ? 0 : (i == Nx - 1 ? 1 : -1);
    const int py = j == 0 ? 2 : (j == Ny - 1 ? 3 : -1);
    const int pz = k == 0 ? 4 : (k == Nz - 1 ? 5 : -1);
    if (px >= 0) v = static_cast<T>(bc.sign[px]) * v + static_cast<T>(bc.value[px]);
    if (py >= 0) v = static_cast<T>(bc.sign[py]) * v + static_cast<T>(bc.value[py]);
    if (pz >= 0) v = static_cast<T>(bc.sign[pz]) * v + static_cast<T>(bc.value[pz]);
    return v;
}

__device__ inline bool written(const BCView &V, long long c) { return (V.wbits[c >> 5] >> (c & 31)) & 1u; }

// Work item t of an idempotent fill, writing into dst: face writes
// (t < n_faces), then ghosts whose chain starts on a cell the fill does not
// write, then ghosts whose chain starts on a face cell.  val(c) must return
// the value cell c has once the producing operator is done, for any cell c
// the fill does not write; every item reads only such cells, so all items
// are independent and the result equals the reference's sequential fill.
// Face ghosts (exactly one coordinate on the ghost shell, so one patch in
// the chain) whose chain starts on an inner fluid cell: with XGrid the
// thread of that cell writes them (face_ghosts) and the XGrid form of
// bc_item takes the remaining ghosts from the view's rest list, so a sweep's
// fill leaves only edges, corners, obstacle faces and their shadows as items.
template <typename T>
__device__ inline void face_ghosts(const LevelDesc &L, const BCView &V, T *dst, long long idx, int i, int j, int k,
                                   T v) {
    if (V.n_ghosts == 0) return;
    // one test for the common case of a cell away from every face (j and k
    // are uniform over a wavefront of the sweep kernels, i is not)
    if (i != V.bc.target[0] && i != V.bc.target[1] && j != V.bc.target[2] && j != V.bc.target[3] &&
        k != V.bc.target[4] && k != V.bc.target[5])
        return;
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    if (i == V.bc.target[0]) dst[idx - i] = static_cast<T>(V.bc.sign[0]) * v + static_cast<T>(V.bc.value[0]);
    if (i == V.bc.target[1])
        dst[idx - i + L.N[0] - 1] = static_cast<T>(V.bc.sign[1]) * v + static_cast<T>(V.bc.value[1]);
    if (j == V.bc.target[2]) dst[idx - j * nj] = static_cast<T>(V.bc.sign[2]) * v + static_cast<T>(V.bc.value[2]);
    if (j == V.bc.target[3])
        dst[idx + (L.N[1] - 1 - j) * nj] = static_cast<T>(V.bc.sign[3]) * v + static_cast<T>(V.bc.value[3]);
    if (k == V.bc.target[4]) dst[idx - k * nk] = static_cast<T>(V.bc.sign[4]) * v + static_cast<T>(V.bc.value[4]);
    if (k == V.bc.target[5])
        dst[idx + (L.N[2] - 1 - k) * nk] = static_cast<T>(V.bc.sign[5]) * v + static_cast<T>(V.bc.value[5]);
}
template <typename T, typename Val, bool XGrid = false>
__device__ inline void bc_item(const LevelDesc &L, const BCView &V, long long t, T *dst, const Val &val) {
    if (t < V.n_faces) {
        const FaceEntry e = V.faces[t];
        const FaceParam q = V.params[e.param];
        if (q.active)
            dst[e.idx] = static_cast<T>(q.sign) * val(static_cast<long long>(e.idx) + e.ref) + static_cast<T>(q.value);
        return;
    }
    t -= V.n_faces;
    long long i, j, k;
    if (XGrid) {
        // the main grid writes the face ghosts of inner fluid cells
        if (t < V.n_rest) {
            const unsigned g = V.rest[t], Nx = static_cast<unsigned>(L.N[0]), Ny = static_cast<unsigned>(L.N[1]);
            i = g % Nx;
            j = (g / Nx) % Ny;
            k = g / (Nx * Ny);
            dst[g] = ghost_chain<T>(L, V.bc, i, j, k, val(ghost_source(L, V.bc, i, j, k)));
            return;
        }
        t -= V.n_rest;
    } else {
        if (t < V.n_ghosts) {
            if (!ghost_coords(L, t, i, j, k)) return;
            const long long s = ghost_source(L, V.bc, i, j, k);
            if (written(V, s)) return;  // produced by a dependent item
            dst[i + static_cast<long long>(L.N[0]) * (j + static_cast<long long>(L.N[1]) * k)] =
                ghost_chain<T>(L, V.bc, i, j, k, val(s));
            return;
        }
        t -= V.n_ghosts;
    }
    if (t < V.n_deps) {
        const DepGhost d = V.deps[t];
        const FaceEntry e = V.faces[d.face];
        const FaceParam q = V.params[e.param];
        const T fv = static_cast<T>(q.sign) * val(static_cast<long long>(e.idx) + e.ref) + static_cast<T>(q.value);
        const long long g = d.idx, Nx = L.N[0], Ny = L.N[1];
        k = g / (Nx * Ny);
        j = (g / Nx) % Ny;
        i = g % Nx;
        dst[g] = ghost_chain<T>(L, V.bc, i, j, k, fv);
    }
}
__device__ __host__ inline long long bc_items(const BCView &V) { return V.n_faces + V.n_ghosts + V.n_deps; }
// items of the XGrid form (the main grid writes the other face ghosts)
__device__ __host__ inline long long bc_items_grid(const BCView &V) { return V.n_faces + V.n_rest + V.n_deps; }

// DomainBoundary.cpp:25-49 for all six patches at once, ghost cell number t
// (0 <= t < ghost_count).  The final value of a ghost cell is the composition
// of the affine maps of the patches containing it, applied in patch order to
// the inner cell the chain of reference offsets ends on.
__device__ inline void domain_ghost(const LevelDesc &L, const DomainBC &bc, double *f, long long t) {
    long long i, j, k;
    if (!ghost_coords(L, t, i, j, k)) return;
    f[lin(L, i, j, k)] = ghost_chain(L, bc, i, j, k, f[ghost_source(L, bc, i, j, k)]);
}

}  // namespace dev
}  // namespace artss
