// HIP kernels of the ARTSS solver hot path for gfx950 (MI355X).
//
// All stencils are HBM-bandwidth bound FP64 work (0.1-0.5 flop/byte), so the
// design goal is one coalesced pass per operator: a workgroup is one 64-wide
// wavefront along x times 4 rows in y, a wavefront reads 512 contiguous bytes
// per array, and inner/obstacle membership is decided from obstacle boxes in
// kernel arguments instead of 8-byte index lists.
//
// Bit-for-bit agreement with the reference serial build: this file is built
// with -ffp-contract=off and every expression keeps the reference's operand
// order (cited per kernel); reductions write per-workgroup partial sums that
// a single workgroup adds in a fixed order, so results are deterministic run
// to run (their rounding differs from the serial left-to-right sum).
#include "kernels.h"
#include "domain.h"
#include "stencil.h"

#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <limits>
#include <mutex>
#include <stdexcept>

namespace artss {
namespace k {

namespace {

constexpr int BX = 64;  // one wavefront along x
constexpr int BY = 4;   // 4 wavefronts per workgroup

inline dim3 inner_grid(const LevelDesc &L) {
    return dim3((L.n[0] + BX - 1) / BX, (L.n[1] + BY - 1) / BY, L.n[2]);
}
inline dim3 full_grid(const LevelDesc &L) {
    return dim3((L.N[0] + BX - 1) / BX, (L.N[1] + BY - 1) / BY, L.N[2]);
}
inline dim3 block() { return dim3(BX, BY, 1); }
// The fused stencil kernels give each thread KZ consecutive cells along z:
// with one cell per thread their waves live about a thousand cycles and the
// workgroup launch rate, not memory, bounds a level-0 launch.
// Small levels keep one cell per thread (their launches are short of waves
// as it is); the count is a kernel argument.
constexpr int KZ = 4;
constexpr long long kBatchCells = 1LL << 20;  // levels with at least this many cells batch
// cells per thread along z chosen by tune_level per level shape and launch
// extent (inner box or whole array; guarded by kz_lock), else the size rule
// above; set_forced_kz overrides both (the values never depend on it)
using TuneKey = std::array<long long, 4>;  // N0, N1, N2, shell_filled
inline TuneKey tune_key(const LevelDesc &L, const BCView &V) {
    return {L.N[0], L.N[1], L.N[2], V.shell_filled ? 1 : 0};
}
std::mutex kz_lock;
std::map<TuneKey, int> kz_tuned;
std::atomic<int> kz_forced{0};
inline int kz_of(const LevelDesc &L, const BCView &V) {
    const int forced = kz_forced.load(std::memory_order_relaxed);
    if (forced > 0) return forced;
    {
        std::lock_guard<std::mutex> g(kz_lock);
        const auto it = kz_tuned.find(tune_key(L, V));
        if (it != kz_tuned.end()) return it->second;
    }
    return L.size >= kBatchCells ? KZ : 1;
}
inline dim3 zbatch(dim3 g, int kz) { return dim3(g.x, g.y, (g.z + kz - 1) / kz); }
inline unsigned grid1d(long long n, int threads) {
    long long g = (n + threads - 1) / threads;
    return static_cast<unsigned>(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

// Block coordinates in the main part of a grid (blockIdx.z < main_z).  The
// hardware deals workgroups to the 8 XCDs round robin by linear id; the
// bijection id -> (id % 8) * per + id / 8 on the first 8 * per ids gives each
// XCD a contiguous run of (x, y, z) blocks, so the z-neighbour planes a
// stencil reads are mostly in that XCD's own L2 instead of being fetched
// again by another XCD.  Cell arithmetic is unchanged; partial sums stay
// indexed by the linear id.
__device__ inline void main_block(int main_z, int &bx, int &by, int &bz) {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned per = gx * gy * static_cast<unsigned>(main_z) / 8u;
    const unsigned nid = id < per * 8u ? (id % 8u) * per + id / 8u : id;
    bx = static_cast<int>(nid % gx);
    by = static_cast<int>((nid / gx) % gy);
    bz = static_cast<int>(nid / (gx * gy));
}

// inner cell of a grid without item blocks
__device__ inline bool inner_cell(const LevelDesc &L, int &i, int &j, int &k) {
    int bx, by, bz;
    main_block(static_cast<int>(gridDim.z), bx, by, bz);
    i = 1 + bx * BX + static_cast<int>(threadIdx.x);
    j = 1 + by * BY + static_cast<int>(threadIdx.y);
    k = 1 + bz;
    return i <= L.n[0] && j <= L.n[1] && !in_obstacle(L, i, j, k);
}

using dev::lin;

// workgroup sum of one double per thread (64x4 block), result valid in thread (0,0)
__device__ inline double block_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __shared__ double wsum[BY];
    if (threadIdx.x == 0) wsum[threadIdx.y] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0 && threadIdx.y == 0)
        for (int w = 0; w < BY; w++) r += wsum[w];
    return r;
}

__device__ inline void write_partial(double *partial, double v) {
    double s = block_sum(v);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
}

// ------------------------------------------------------------------ element-wise
__global__ void k_copy(double *__restrict__ d, const double *__restrict__ s, long long n, const int *gate) {
    if (gate != nullptr && *gate) return;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] = s[i];
}
// each source into two destinations, for F (1..4) fields at once
template <int F>
__global__ void k_copy_fan(CopyFan f, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
#pragma unroll
        for (int q = 0; q < F; q++) {
            const double v = f.src[q][i];
            f.dst0[q][i] = v;
            f.dst1[q][i] = v;
        }
    }
}
__global__ void k_fill32(float *__restrict__ d, float v, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] = v;
}
__global__ void k_fill(double *__restrict__ d, double v, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] = v;
}
__global__ void k_scale(double *__restrict__ d, double v, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] *= v;
}
__global__ void k_add(double *__restrict__ d, const double *__restrict__ s, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] += s[i];
}
__global__ void k_scale_to(double *__restrict__ d, const double *__restrict__ s, double v, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] = s[i] * v;
}
__global__ void k_copy_many(double *const *__restrict__ dst, const double *const *__restrict__ src, long long n) {
    double *d = dst[blockIdx.y];
    const double *s = src[blockIdx.y];
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] = s[i];
}
__global__ void k_scale_to_dev(double *__restrict__ d, const double *__restrict__ s, const double *__restrict__ v,
                               long long n) {
    const double f = *v;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) d[i] = s[i] * f;
}

// copy over inner + ghost shell, leaving obstacle cells alone
__global__ void k_copy_domain(LevelDesc L, double *__restrict__ d, const double *__restrict__ s) {
    const int i = static_cast<int>(blockIdx.x) * BX + static_cast<int>(threadIdx.x);
    const int j = static_cast<int>(blockIdx.y) * BY + static_cast<int>(threadIdx.y);
    const int k = static_cast<int>(blockIdx.z);
    if (i >= L.N[0] || j >= L.N[1]) return;
    const bool ghost = i == 0 || j == 0 || k == 0 || i == L.N[0] - 1 || j == L.N[1] - 1 || k == L.N[2] - 1;
    if (!ghost && in_obstacle(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    d[idx] = s[idx];
}

// ------------------------------------------------------------------ boundary
// ObstacleBoundary.cpp:42-48: f[idx] = sign * f[idx +- ref] + value
__global__ void k_bc_faces(const FaceEntry *__restrict__ e, long long n, const FaceParam *__restrict__ p,
                           double *__restrict__ f, const int *gate) {
    if (gate != nullptr && *gate) return;
    const long long t = blockIdx.x * 256LL + threadIdx.x;
    if (t >= n) return;
    dev::face_write(e[t], p, f);
}

// DomainBoundary.cpp:25-49 for all six patches at once.  Each ghost cell is
// owned by one thread; its final value is the composition of the affine maps
// of the patches containing it, applied in patch order to the inner cell the
// chain of reference offsets ends on.
__global__ void k_bc_domain(LevelDesc L, DomainBC bc, double *__restrict__ f, const int *gate) {
    if (gate != nullptr && *gate) return;
    dev::domain_ghost(L, bc, f, blockIdx.x * 256LL + threadIdx.x);
}

// Obstacle faces and domain ghosts in one launch, for a boundary whose
// writes read no cell that any of its writes changes (so faces and ghosts
// can run in any order and the fill is idempotent).
__global__ void k_bc_merged(LevelDesc L, DomainBC bc, const FaceEntry *__restrict__ e, long long n_faces,
                            const FaceParam *__restrict__ p, double *__restrict__ f) {
    const long long t = blockIdx.x * 256LL + threadIdx.x;
    if (t < n_faces) {
        dev::face_write(e[t], p, f);
        return;
    }
    if (bc.active) dev::domain_ghost(L, bc, f, t - n_faces);
}

// ------------------------------------------------------------------ advection
// SLAdvect.cpp:71-163.  fmod(C, 1) of a finite C > 0 equals C - trunc(C)
// exactly (the difference is representable, Sterbenz), so no library call.
__global__ void k_sl_advect(LevelDesc L, double dtx, double dty, double dtz, double *__restrict__ out,
                            const double *__restrict__ in, const double *__restrict__ u,
                            const double *__restrict__ v, const double *__restrict__ w) {
    int ci, cj, ck;
    if (!inner_cell(L, ci, cj, ck)) return;
    const long long idx = lin(L, ci, cj, ck);
    const double eps = 1e-6;
    const long i = ci, j = cj, k = ck;
    const double Ci = dtx * u[idx], Cj = dty * v[idx], Ck = dtz * w[idx];
    long i0 = i, i1 = i, j0 = j, j1 = j, k0 = k, k1 = k;
    double r = 1, s = 1, t = 1;
    // cell shift (long)C; |C| is capped at 1e12 only so the cast cannot
    // overflow (the clamp to the inner range below already covers any such C)
    auto shift = [](double C) { return static_cast<long>(fmin(fmax(C, -1e12), 1e12)); };
    if (Ci > eps) {
        i0 = max(1L, i - shift(Ci));
        i1 = i0 - 1;
        r = Ci - trunc(Ci);
    } else if (Ci < -eps) {
        i0 = min(static_cast<long>(L.n[0]), i - shift(Ci));
        i1 = i0 + 1;
        r = (-Ci) - trunc(-Ci);
    }
    if (Cj > eps) {
        j0 = max(1L, j - shift(Cj));
        j1 = j0 - 1;
        s = Cj - trunc(Cj);
    } else if (Cj < -eps) {
        j0 = min(static_cast<long>(L.n[1]), j - shift(Cj));
        j1 = j0 + 1;
        s = (-Cj) - trunc(-Cj);
    }
    if (Ck > eps) {
        k0 = max(1L, k - shift(Ck));
        k1 = k0 - 1;
        t = Ck - trunc(Ck);
    } else if (Ck < -eps) {
        k0 = min(static_cast<long>(L.n[2]), k - shift(Ck));
        k1 = k0 + 1;
        t = (-Ck) - trunc(-Ck);
    }
    const double d000 = in[lin(L, i0, j0, k0)], d100 = in[lin(L, i1, j0, k0)];
    const double d010 = in[lin(L, i0, j1, k0)], d110 = in[lin(L, i1, j1, k0)];
    const double d001 = in[lin(L, i0, j0, k1)], d101 = in[lin(L, i1, j0, k1)];
    const double d011 = in[lin(L, i0, j1, k1)], d111 = in[lin(L, i1, j1, k1)];
    const double r100 = d000 + r * (d100 - d000);
    const double r110 = d010 + r * (d110 - d010);
    const double r101 = d001 + r * (d101 - d001);
    const double r111 = d011 + r * (d111 - d011);
    const double s110 = r100 + s * (r110 - r100);
    const double s111 = r101 + s * (r111 - r101);
    out[idx] = s110 + t * (s111 - s110);
}

// ------------------------------------------------------------------ Jacobi
struct DiffCoef {
    double ax, ay, az, rb;   // constant form
    double D, dt, rdx, rdy, rdz;
};

// JacobiDiffuse.cpp:241-249 (ev == nullptr) and :345-362 (turbulent), plus the
// update-size norm of :72-77 / :162-172 as per-workgroup partial sums.
__global__ void k_jacobi_diffuse(LevelDesc L, DiffCoef c, double w, double *__restrict__ out,
                                 const double *__restrict__ in, const double *__restrict__ b,
                                 const double *__restrict__ ev, double *__restrict__ partial) {
    int i, j, k;
    double sq = 0.0;
    if (inner_cell(L, i, j, k)) {
        const long long idx = lin(L, i, j, k);
        const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
        const double dsign = 1;
        double rb, out_h;
        if (ev == nullptr) {
            rb = c.rb;
            out_h = (dsign * b[idx] + c.ax * (in[idx + 1] + in[idx - 1]) + c.ay * (in[idx + nj] + in[idx - nj]) +
                     c.az * (in[idx + nk] + in[idx - nk])) / rb;
        } else {
            const double e = ev[idx];
            const double aX = (c.D + e) * c.dt * c.rdx * c.rdx;
            const double aY = (c.D + e) * c.dt * c.rdy * c.rdy;
            const double aZ = (c.D + e) * c.dt * c.rdz * c.rdz;
            rb = (1. + 2. * (aX + aY + aZ));
            const double bb = 1. / rb;
            const double d_aX = aX * (in[idx + 1] + in[idx - 1]);
            const double d_aY = aY * (in[idx + nj] + in[idx - nj]);
            const double d_aZ = aZ * (in[idx + nk] + in[idx - nk]);
            out_h = bb * (dsign * b[idx] + d_aX + d_aY + d_aZ);
        }
        const double x = in[idx];
        const double o = (1 - w) * x + w * out_h;
        out[idx] = o;
        const double res = rb * (o - x);
        sq = res * res;
    }
    write_partial(partial, sq);
}

// JacobiDiffuse.cpp:292-300
__global__ void k_jacobi_mg(LevelDesc L, double ax, double ay, double az, double beta, double dsign, double w,
                            double *__restrict__ out, const double *__restrict__ in, const double *__restrict__ b) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    out[idx] = dev::jacobi_mg(L, ax, ay, az, beta, dsign, w, in, b, idx);
}

// First sweep of a Jacobi smoothing with the smoother's initial copy folded
// in (JacobiDiffuse.cpp:85-99 + :292-300): every cell of dst is written,
// inner cells with the update of src, all others with src itself.
__global__ void k_jacobi_mg_first(LevelDesc L, double ax, double ay, double az, double beta, double dsign,
                                  double w, double *__restrict__ dst, const double *__restrict__ src,
                                  const double *__restrict__ b) {
    int bx, by, bz;
    main_block(static_cast<int>(gridDim.z), bx, by, bz);
    const int i = bx * BX + static_cast<int>(threadIdx.x);
    const int j = by * BY + static_cast<int>(threadIdx.y);
    const int k = bz;
    if (i >= L.N[0] || j >= L.N[1]) return;
    const long long idx = lin(L, i, j, k);
    const bool inner = i >= 1 && j >= 1 && k >= 1 && i <= L.n[0] && j <= L.n[1] && k <= L.n[2] &&
                       !in_obstacle(L, i, j, k);
    dst[idx] = inner ? dev::jacobi_mg(L, ax, ay, az, beta, dsign, w, src, b, idx) : src[idx];
}

// ------------------------------------------------------------------ explicit diffusion
// ExplicitDiffuse.cpp:49-78 (ev == nullptr) and :81-121: one explicit Euler
// step on inner cells; the turbulent form has no 1/h^2 and subtracts the
// upper flux from the lower one, both as the reference writes them.
__global__ void k_explicit_diffuse(LevelDesc L, double rdx, double rdy, double rdz, double D, double dt,
                                   double *__restrict__ out, const double *__restrict__ in,
                                   const double *__restrict__ ev) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long ni = 1, nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    if (ev == nullptr) {
        out[idx] = in[idx] + dt * ((in[idx + ni] - 2 * in[idx] + in[idx - ni]) * rdx +
                                   (in[idx + nj] - 2 * in[idx] + in[idx - nj]) * rdy +
                                   (in[idx + nk] - 2 * in[idx] + in[idx - nk]) * rdz);
        return;
    }
    const double nu_x1 = 0.5 * (ev[idx + ni] + ev[idx]) + D;
    const double nu_x2 = 0.5 * (ev[idx - ni] + ev[idx]) + D;
    const double nu_y1 = 0.5 * (ev[idx + nj] + ev[idx]) + D;
    const double nu_y2 = 0.5 * (ev[idx - nj] + ev[idx]) + D;
    const double nu_z1 = 0.5 * (ev[idx + nk] + ev[idx]) + D;
    const double nu_z2 = 0.5 * (ev[idx - nk] + ev[idx]) + D;
    const double di0 = (in[idx] - in[idx - ni]);
    const double di1 = (in[idx + ni] - in[idx]);
    const double dj0 = (in[idx] - in[idx - nj]);
    const double dj1 = (in[idx + nj] - in[idx]);
    const double dk0 = (in[idx] - in[idx - nk]);
    const double dk1 = (in[idx + nk] - in[idx]);
    out[idx] = in[idx] + dt * ((nu_x1 * di0 - nu_x2 * di1) + (nu_y1 * dj0 - nu_y2 * dj1) + (nu_z1 * dk0 - nu_z2 * dk1));
}

// ------------------------------------------------------------------ dissipation
// ISource.cpp:85-114: the expression as the reference writes it (every
// difference quotient formed the same way each time it appears)
__global__ void k_dissipate(LevelDesc L, double rdx, double rdy, double rdz, double nu, double dt,
                            double *__restrict__ out, const double *__restrict__ in_u,
                            const double *__restrict__ in_v, const double *__restrict__ in_w) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long ni = 1, nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const double ux = 0.5 * rdx * (in_u[idx + ni] - in_u[idx - ni]);
    const double vy = 0.5 * rdy * (in_v[idx + nj] - in_v[idx - nj]);
    const double wz = 0.5 * rdz * (in_w[idx + nk] - in_w[idx - nk]);
    const double vx = 0.5 * rdx * (in_v[idx + ni] - in_v[idx - ni]);
    const double uy = 0.5 * rdy * (in_u[idx + nj] - in_u[idx - nj]);
    const double wy = 0.5 * rdy * (in_w[idx + nj] - in_w[idx - nj]);
    const double vz = 0.5 * rdz * (in_v[idx + nk] - in_v[idx - nk]);
    const double uz = 0.5 * rdz * (in_u[idx + nk] - in_u[idx - nk]);
    const double wx = 0.5 * rdx * (in_w[idx + ni] - in_w[idx - ni]);
    const double out_h = nu * (2 * ux * ux + 2 * vy * vy + 2 * wz * wz + (vx + uy) * (vx + uy) +
                               (wy + vz) * (wy + vz) + (uz + wx) * (uz + wx));
    out[idx] += dt * out_h;
}

// ------------------------------------------------------------------ colored Gauss-Seidel
// One color of ColoredGaussSeidelDiffuse::colored_gauss_seidel_step
// (:212-222, :270-294) with the stencil of :316-341.  Cells of one color only
// read cells of the other color, so the order inside a color is free.
// Thread t of a row handles i = i_first + 2t.
__global__ void k_cgs_color(LevelDesc L, int color, double ax, double ay, double az, double beta, double dsign,
                            double w, double D, const double *__restrict__ ev, double dt, double rdx, double rdy,
                            double rdz, double *__restrict__ out, const double *__restrict__ b, const int *gate) {
    if (gate != nullptr && *gate) return;
    int bx, by, bz;
    main_block(static_cast<int>(gridDim.z), bx, by, bz);
    const int t = bx * BX + static_cast<int>(threadIdx.x);
    const int j = 1 + by * BY + static_cast<int>(threadIdx.y);
    const int k = 1 + bz;
    if (j > L.n[1]) return;
    const int i = (((1 + j + k) & 1) == color ? 1 : 2) + 2 * t;
    if (i > L.n[0] || in_obstacle(L, i, j, k)) return;
    const long long index = lin(L, i, j, k);
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    double aX = ax, aY = ay, aZ = az, bb = beta;
    if (ev != nullptr) {
        const double e = ev[index];
        aX = (D + e) * dt * rdx * rdx;
        aY = (D + e) * dt * rdy * rdy;
        aZ = (D + e) * dt * rdz * rdz;
        const double rb = (1. + 2. * (aX + aY + aZ));
        bb = 1. / rb;
    }
    const double ox = out[index + 1], ox2 = out[index - 1];
    const double oy = out[index + nj], oy2 = out[index - nj];
    const double oz = out[index + nk], oz2 = out[index - nk];
    const double db = b[index], r_out = out[index];
    const double out_h = bb * (dsign * db + aX * (ox + ox2) + aY * (oy + oy2) + aZ * (oz + oz2));
    out[index] = (1 - w) * r_out + w * out_h;
}

__global__ void k_update_norm(LevelDesc L, DiffCoef c, const double *__restrict__ out, const double *__restrict__ in,
                              const double *__restrict__ ev, double *__restrict__ partial, const int *gate) {
    if (gate != nullptr && *gate) return;  // uniform over the grid
    int i, j, k;
    double sq = 0.0;
    if (inner_cell(L, i, j, k)) {
        const long long idx = lin(L, i, j, k);
        double rb = c.rb;
        if (ev != nullptr) {
            const double e = ev[idx];
            const double aX = (c.D + e) * c.dt * c.rdx * c.rdx;
            const double aY = (c.D + e) * c.dt * c.rdy * c.rdy;
            const double aZ = (c.D + e) * c.dt * c.rdz * c.rdz;
            rb = (1. + 2. * (aX + aY + aZ));
        }
        const double r = rb * (out[idx] - in[idx]);
        sq = r * r;
    }
    write_partial(partial, sq);
}

// ------------------------------------------------------------------ multigrid
using dev::laplace;

// VCycleMG.cpp:346-354
__global__ void k_residuum(LevelDesc L, double *__restrict__ out, const double *__restrict__ x,
                           const double *__restrict__ b) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const double weighted = laplace(L, x, idx);
    out[idx] = b[idx] - weighted;
}

// VCycleMG.cpp:186-193 / :668-675 / :777-785
__global__ void k_residual_partials(LevelDesc L, const double *__restrict__ x, const double *__restrict__ b,
                                    double *__restrict__ partial) {
    int i, j, k;
    double sq = 0.0;
    if (inner_cell(L, i, j, k)) {
        const long long idx = lin(L, i, j, k);
        const double r = b[idx] - laplace(L, x, idx);
        sq = r * r;
    }
    write_partial(partial, sq);
}

// VCycleMG.cpp:404-421, iterating the coarse inner cells
__global__ void k_restrict(LevelDesc Lc, LevelDesc Lf, double *__restrict__ out, const double *__restrict__ in) {
    int i, j, k;
    if (!inner_cell(Lc, i, j, k)) return;
    out[lin(Lc, i, j, k)] = dev::restrict_cell(Lf, in, i, j, k);
}

// VCycleMG.cpp:466-546: eight fine cells per coarse inner cell, in the
// reference's write order
__global__ void k_prolongate(LevelDesc Lc, LevelDesc Lf, double *__restrict__ out, const double *__restrict__ in) {
    int i, j, k;
    if (!inner_cell(Lc, i, j, k)) return;
    const long long idx = lin(Lc, i, j, k);
#pragma unroll
    for (int q = 0; q < 8; q++) out[dev::prolong_target(Lf, i, j, k, q)] = dev::prolong_value(Lc, in, idx, q);
}

// Restriction (VCycleMG.cpp:404-421) fused with the reset of the coarse
// error (VCycleMG.cpp:250): one thread per coarse cell of the full array.
template <typename Tin, typename Tout, typename Terr>
__global__ void k_restrict_fill(LevelDesc Lc, LevelDesc Lf, Tout *__restrict__ out, const Tin *__restrict__ in,
                                Terr *__restrict__ err) {
    const int i = static_cast<int>(blockIdx.x) * BX + static_cast<int>(threadIdx.x);
    const int j = static_cast<int>(blockIdx.y) * BY + static_cast<int>(threadIdx.y);
    const int k = static_cast<int>(blockIdx.z);
    if (i >= Lc.N[0] || j >= Lc.N[1]) return;
    const long long idx = lin(Lc, i, j, k);
    err[idx] = 0;
    if (i >= 1 && j >= 1 && k >= 1 && i <= Lc.n[0] && j <= Lc.n[1] && k <= Lc.n[2] && !in_obstacle(Lc, i, j, k))
        out[idx] = static_cast<Tout>(dev::restrict_t<Tin>(Lf, in, i, j, k));
}

// Prolongation added straight into the fine error (VCycleMG.cpp:266-270
// without the intermediate field), for boundaries whose fill overwrites
// every non-inner cell the addition would touch.
__global__ void k_prolongate_add(LevelDesc Lc, LevelDesc Lf, double *__restrict__ out, const double *__restrict__ in) {
    int i, j, k;
    if (!inner_cell(Lc, i, j, k)) return;
    const long long idx = lin(Lc, i, j, k);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const long long t = dev::prolong_target(Lf, i, j, k, q);
        out[t] = out[t] + dev::prolong_value(Lc, in, idx, q);
    }
}

// ------------------------------------------------------------------ fused V-cycle kernels
// Kernels that finish an operator and the boundary fill that follows it in
// one launch.  They need an idempotent fill (BCView): every filled value is
// a function of cells the fill does not write, so the fill's work items run
// in extra workgroups of the same launch, each re-evaluating the operator at
// the one cell it reads (the same expression, hence the same bits).

// The extra workgroups are appended as additional z-slices of the grid.
inline dim3 with_items(dim3 g, long long items) {
    const long long per_slice = static_cast<long long>(g.x) * g.y * BX * BY;
    return dim3(g.x, g.y, g.z + static_cast<unsigned>((items + per_slice - 1) / per_slice));
}
__device__ inline long long item_index(int main_z) {
    return ((static_cast<long long>(blockIdx.z - main_z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) *
               (BX * BY) + threadIdx.y * BX + threadIdx.x;
}
// (32-bit arithmetic: a level has fewer than 2^32 cells)
__device__ inline bool is_inner(const LevelDesc &L, long long c, int &i, int &j, int &k) {
    const unsigned u = static_cast<unsigned>(c), Nx = static_cast<unsigned>(L.N[0]), Ny = static_cast<unsigned>(L.N[1]);
    k = static_cast<int>(u / (Nx * Ny));
    j = static_cast<int>((u / Nx) % Ny);
    i = static_cast<int>(u % Nx);
    return i >= 1 && j >= 1 && k >= 1 && i <= L.n[0] && j <= L.n[1] && k <= L.n[2] && !in_obstacle(L, i, j, k);
}

// ---- colored Gauss-Seidel iteration
// ColoredGaussSeidelDiffuse.cpp:111-173 (the turbulent form; ev == nullptr
// gives the constant coefficients of :38-98).  Every update is the
// expression of k_cgs_color, so the values are those of the reference's
// in-place sweeps and fill.
// the update with the eddy viscosity e at the cell (turb) or the constant
// coefficients
__device__ __forceinline__ double cgs_value_e(const CgsCoef &c, bool turb, double e, double r_out, double ox,
                                              double ox2, double oy, double oy2, double oz, double oz2, double db,
                                              double *rb_out) {
    double aX = c.ax, aY = c.ay, aZ = c.az, bb = c.beta, rb = c.rb;
    if (turb) {
        aX = (c.D + e) * c.dt * c.rdx * c.rdx;
        aY = (c.D + e) * c.dt * c.rdy * c.rdy;
        aZ = (c.D + e) * c.dt * c.rdz * c.rdz;
        rb = (1. + 2. * (aX + aY + aZ));
        bb = 1. / rb;
    }
    if (rb_out) *rb_out = rb;
    const double out_h = bb * (c.dsign * db + aX * (ox + ox2) + aY * (oy + oy2) + aZ * (oz + oz2));
    return (1 - c.w) * r_out + c.w * out_h;
}
__device__ __forceinline__ double cgs_value(const CgsCoef &c, const double *__restrict__ ev, long long index,
                                            double r_out, double ox, double ox2, double oy, double oy2, double oz,
                                            double oz2, double db, double *rb_out) {
    return cgs_value_e(c, ev != nullptr, ev != nullptr ? ev[index] : 0.0, r_out, ox, ox2, oy, oy2, oz, oz2, db,
                       rb_out);
}
// the update of fluid cell idx: neighbours from nb, its own old value from
// old (the two arrays may be the same)
__device__ __forceinline__ double cgs_update(const LevelDesc &L, const CgsCoef &c, const double *__restrict__ ev,
                                             const double *nb, const double *old, const double *__restrict__ b,
                                             long long idx, double *rb_out) {
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    return cgs_value(c, ev, idx, old[idx], nb[idx + 1], nb[idx - 1], nb[idx + nj], nb[idx - nj], nb[idx + nk],
                     nb[idx - nk], b[idx], rb_out);
}

// ---- one colored Gauss-Seidel iteration in one launch
// Out of place: src holds the iterate with its fill (what the reference's
// `in` holds after its copy), dst receives the next iterate on every cell
// (the fluid cells both colours' updates, the cells the fill writes the
// fill, every other cell src's value), and nothing is read from dst.  The
// first colour is the fluid cells with i + j + k even; a second-colour cell
// reads only first-colour neighbours, with the values they hold after the
// first colour, and cells off the fluid as src holds them (for those the
// fill writes: the previous fill).
//
// A workgroup of kCgsWaves wavefronts owns a tile of kCgsOwnX x kCgsOwnY
// columns (i, j) and marches along z through kCgsPlanes planes.  Each lane
// holds two cells adjacent along x, so that every lane updates one cell of
// each colour per plane; a wavefront covers two rows of 64 cells, and the
// tile has a one-cell halo of columns and rows.  At step t each lane
// evaluates the first colour at plane t + 1 (src from memory) and puts the
// values its two cells hold after the first colour into a three-plane ring
// in LDS; after one barrier it updates the second colour at plane t from
// the ring (x, y) and its own registers (z), writes dst, applies the fill
// through its face ghosts and adds the update norm.  Fill items run in
// extra workgroups that re-evaluate both colours from src at the cell each
// write reads.
//
// Memory is read through buffer descriptors with 32-bit offsets: a load
// outside an array returns 0 instead of touching memory, so every lane
// issues the same loads at every step, whatever cells it holds, with no
// branch around them (a value read outside its array reaches no update:
// the neighbours of a fluid cell lie inside the array).
// (4 wavefronts and 16 planes measured fastest on the 256x128x128 plume:
// 71.5 us per iteration against 73.9 for 8 wavefronts, 81.9 for 8 planes)
constexpr int kCgsWaves = 4;
constexpr int kCgsRows = 2 * kCgsWaves;  // rows of a tile with its halo
constexpr int kCgsOwnX = 62;             // of 64 cells per row
constexpr int kCgsOwnY = kCgsRows - 2;
constexpr int kCgsPlanes = 16;

struct CgsTiles {
    int tx, ty, tz;  // tiles along x, y and z; main workgroups tx * ty * tz
};

__device__ __forceinline__ bool fluid_at(const LevelDesc &L, int i, int j, int k) {
    return i >= 1 && j >= 1 && k >= 1 && i <= L.n[0] && j <= L.n[1] && k <= L.n[2] && !in_obstacle(L, i, j, k);
}

// the value cell n holds after the first colour of the iteration
__device__ __forceinline__ double cgs_first(const LevelDesc &L, const CgsCoef &c, const double *__restrict__ ev,
                                            const double *__restrict__ b, const double *__restrict__ src,
                                            long long n) {
    int i, j, k;
    if (is_inner(L, n, i, j, k) && ((i + j + k) & 1) == 0) return cgs_update(L, c, ev, src, src, b, n, nullptr);
    return src[n];
}

// raw buffer descriptor of n bytes at p (loads at offsets >= n return 0)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void *p, unsigned n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), static_cast<short>(0), static_cast<int>(n),
                                             0x00020000);
}
__device__ __forceinline__ double buffer_f64(__amdgpu_buffer_rsrc_t r, unsigned offset) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, offset, 0, 0));
}
__device__ __forceinline__ unsigned buffer_u32(__amdgpu_buffer_rsrc_t r, unsigned offset) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, offset, 0, 0);
}

template <bool TURB>
__global__ void __launch_bounds__(64 * kCgsWaves)
    k_cgs_iter(LevelDesc L, BCView V, CgsTiles T, CgsCoef c, const double *__restrict__ ev,
               const double *__restrict__ b, const double *__restrict__ src, double *__restrict__ dst,
               double *__restrict__ partial, const int *gate) {
    if (gate != nullptr && *gate) return;  // uniform over the grid
    const unsigned main_blocks = static_cast<unsigned>(T.tx) * T.ty * T.tz;
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    if (blockIdx.x >= main_blocks) {
        // fill items: the value the iteration gives the cell each write reads
        const long long t = static_cast<long long>(blockIdx.x - main_blocks) * (64 * kCgsWaves) + threadIdx.y * 64 +
                            threadIdx.x;
        auto val = [&](long long q) -> double {
            int i, j, k;
            if (!is_inner(L, q, i, j, k)) return src[q];
            if (((i + j + k) & 1) == 0) return cgs_update(L, c, ev, src, src, b, q, nullptr);
            // the six neighbours after the first colour (one evaluation in a
            // loop keeps the item code small)
            double nb[6];
#pragma unroll 1
            for (int d = 0; d < 6; d++) {
                const long long step = (d >> 1) == 0 ? 1 : ((d >> 1) == 1 ? nj : nk);
                const double x = cgs_first(L, c, ev, b, src, (d & 1) ? q - step : q + step);
                nb[0] = d == 0 ? x : nb[0];
                nb[1] = d == 1 ? x : nb[1];
                nb[2] = d == 2 ? x : nb[2];
                nb[3] = d == 3 ? x : nb[3];
                nb[4] = d == 4 ? x : nb[4];
                nb[5] = d == 5 ? x : nb[5];
            }
            return cgs_value(c, ev, q, src[q], nb[0], nb[1], nb[2], nb[3], nb[4], nb[5], b[q], nullptr);
        };
        dev::bc_item<double, decltype(val), true>(L, V, t, dst, val);
        return;
    }
    // consecutive z tiles, then y tiles, on one XCD (the hardware deals
    // workgroups to the 8 XCDs round robin)
    const unsigned id = blockIdx.x, per = main_blocks / 8u;
    const unsigned nid = id < per * 8u ? (id % 8u) * per + id / 8u : id;
    const int tz = static_cast<int>(nid % T.tz);
    const int ty = static_cast<int>((nid / T.tz) % T.ty);
    const int tx = static_cast<int>(nid / (T.tz * T.ty));
    const int lane = static_cast<int>(threadIdx.x);
    const int r = 2 * static_cast<int>(threadIdx.y) + (lane >> 5);  // row of the tile
    const int pr = lane & 31;                                       // pair of the row
    const int xa = tx * kCgsOwnX - 2 + 2 * pr;                      // the lane's cells xa, xa + 1 (xa even)
    const int j = ty * kCgsOwnY - 1 + r;
    const bool own_row = r >= 1 && r <= kCgsOwnY && j >= 0 && j < L.N[1];
    const bool own_a = own_row && pr >= 1 && xa >= 0 && xa < L.N[0];
    const bool own_b = own_row && pr <= 30 && xa + 1 >= 0 && xa + 1 < L.N[0];
    const int k0 = tz * kCgsPlanes;
    const int k1 = min(k0 + kCgsPlanes, L.N[2]);
    constexpr bool turb = TURB;
    // columns of the two cells: inside the inner box, and inside each obstacle's x-y extent (bit o)
    const bool in_a = xa >= 1 && xa <= L.n[0] && j >= 1 && j <= L.n[1];
    const bool in_b = xa + 1 >= 1 && xa + 1 <= L.n[0] && j >= 1 && j <= L.n[1];
    unsigned ob_a = 0, ob_b = 0;
    for (int o = 0; o < L.n_obst; o++) {
        const Box &q = L.obst[o];
        const bool yin = j >= q.lo[1] && j <= q.hi[1];
        ob_a |= (yin && xa >= q.lo[0] && xa <= q.hi[0]) ? 1u << o : 0u;
        ob_b |= (yin && xa + 1 >= q.lo[0] && xa + 1 <= q.hi[0]) ? 1u << o : 0u;
    }
    // obstacles whose z extent holds plane k (uniform)
    auto zmask = [&](int k) {
        unsigned m = 0;
        for (int o = 0; o < L.n_obst; o++) m |= (k >= L.obst[o].lo[2] && k <= L.obst[o].hi[2]) ? 1u << o : 0u;
        return m;
    };
    const unsigned bytes = static_cast<unsigned>(L.size) * 8u;
    const __amdgpu_buffer_rsrc_t rs = buffer_of(src, bytes), rb_ = buffer_of(b, bytes),
                                 re = buffer_of(TURB ? ev : b, bytes),
                                 rw = buffer_of(V.wbits, static_cast<unsigned>((L.size + 31) / 32) * 4u);
    // byte offset of cell xa of row j in plane 0 (wraps for cells before the array)
    const unsigned lo = static_cast<unsigned>(xa + static_cast<int>(nj) * j) * 8u;
    const unsigned pn = static_cast<unsigned>(nk) * 8u, rn = static_cast<unsigned>(nj) * 8u;
    __shared__ double ring[3][kCgsRows][64];
    // every value step t reads from memory (cells xa, xa + 1; slot s of the
    // pair is the first-colour cell at plane t + 1 and the second-colour
    // cell at plane t, the same one): src of both cells at planes t and t + 1;
    // at the slot's cell src at plane t + 2, its outer x neighbour and its
    // y neighbours at t + 1, b and the eddy viscosity at t + 1, b at t; the
    // eddy viscosity of both cells and the fill's bits at t
    struct Fetch {
        double s0a, s0b, s1a, s1b, s2, xo, yp, ym, bn, en, bc, eca, ecb;
        unsigned wa, wb;
    };
    auto fetch = [&](int t) -> Fetch {
        const unsigned o0 = lo + static_cast<unsigned>(t) * pn, o1 = o0 + pn, o2 = o1 + pn;
        const bool sa = ((j + t) & 1) == 1;  // slot 0 (cell xa)
        const unsigned e = sa ? 0u : 8u;
        const unsigned ca = o0 / 8u;  // cell index of xa at plane t (wrapped when outside)
        Fetch f;
        f.s0a = buffer_f64(rs, o0);
        f.s0b = buffer_f64(rs, o0 + 8u);
        f.s1a = buffer_f64(rs, o1);
        f.s1b = buffer_f64(rs, o1 + 8u);
        f.s2 = buffer_f64(rs, o2 + e);
        f.xo = buffer_f64(rs, sa ? o1 - 8u : o1 + 16u);
        f.yp = buffer_f64(rs, o1 + e + rn);
        f.ym = buffer_f64(rs, o1 + e - rn);
        f.bn = buffer_f64(rb_, o1 + e);
        f.en = TURB ? buffer_f64(re, o1 + e) : 0.0;
        f.bc = buffer_f64(rb_, o0 + e);
        f.eca = TURB ? buffer_f64(re, o0) : 0.0;
        f.ecb = TURB ? buffer_f64(re, o0 + 8u) : 0.0;
        f.wa = buffer_u32(rw, (ca >> 5) * 4u);
        f.wb = buffer_u32(rw, ((ca + 1u) >> 5) * 4u);
        return f;
    };
    // the two cells' values after the first colour at planes t - 1 and t
    double pa = 0.0, pb = 0.0, ma = 0.0, mb = 0.0;
    double sq = 0.0;
    auto rb_of = [&](double e) {
        return turb ? (1. + 2. * ((c.D + e) * c.dt * c.rdx * c.rdx + (c.D + e) * c.dt * c.rdy * c.rdy +
                                  (c.D + e) * c.dt * c.rdz * c.rdz))
                    : c.rb;
    };
    // step t with the loads fetched for it
    auto step = [&](int t, const Fetch &f) {
        // first colour at plane t + 1: the slot's cell
        const int kn = t + 1;
        const bool kin = kn >= 1 && kn <= L.n[2];
        const unsigned zm = zmask(kn);
        const bool sa = ((j + t) & 1) == 1;  // slot 0: xa + j + t + 1 even
        const bool fl = sa ? (in_a && kin && (ob_a & zm) == 0) : (in_b && kin && (ob_b & zm) == 0);
        double na = f.s1a, nb = f.s1b;
        if (fl) {
            const double v = cgs_value_e(c, turb, f.en, sa ? f.s1a : f.s1b, sa ? f.s1b : f.xo, sa ? f.xo : f.s1a,
                                         f.yp, f.ym, f.s2, sa ? f.s0a : f.s0b, f.bn, nullptr);
            na = sa ? v : na;
            nb = sa ? nb : v;
        }
        double *slot = &ring[(t + 4) % 3][r][2 * pr];
        slot[0] = na;
        slot[1] = nb;
        __syncthreads();
        // second colour, fill and norm at plane t: the slot's cell
        if (t >= k0 && t < k1 && (own_a || own_b)) {
            const bool tin = t >= 1 && t <= L.n[2];
            const unsigned zc = zmask(t);
            const bool fa = own_a && in_a && tin && (ob_a & zc) == 0;
            const bool fb = own_b && in_b && tin && (ob_b & zc) == 0;
            const int q = (t + 3) % 3;
            double fin_a = ma, fin_b = mb, rba, rbb;
            if (sa ? fa : fb) {
                const int x = 2 * pr + (sa ? 0 : 1);  // column of the tile
                double rbs;
                const double v = cgs_value_e(c, turb, sa ? f.eca : f.ecb, sa ? f.s0a : f.s0b, ring[q][r][x + 1],
                                             ring[q][r][x - 1], ring[q][r + 1][x], ring[q][r - 1][x],
                                             sa ? na : nb, sa ? pa : pb, f.bc, &rbs);
                fin_a = sa ? v : fin_a;
                fin_b = sa ? fin_b : v;
                rba = sa ? rbs : rb_of(f.eca);
                rbb = sa ? rb_of(f.ecb) : rbs;
            } else {
                rba = rb_of(f.eca);
                rbb = rb_of(f.ecb);
            }
            const long long ia = static_cast<long long>(xa) + nj * j + nk * t;
            if (fa) {
                dst[ia] = fin_a;
                dev::face_ghosts(L, V, dst, ia, xa, j, t, fin_a);
                const double d = rba * (fin_a - f.s0a);
                sq += d * d;
            } else if (own_a && ((f.wa >> (static_cast<unsigned>(ia) & 31u)) & 1u) == 0) {
                dst[ia] = f.s0a;  // not written by the fill
            }
            if (fb) {
                dst[ia + 1] = fin_b;
                dev::face_ghosts(L, V, dst, ia + 1, xa + 1, j, t, fin_b);
                const double d = rbb * (fin_b - f.s0b);
                sq += d * d;
            } else if (own_b && ((f.wb >> (static_cast<unsigned>(ia + 1) & 31u)) & 1u) == 0) {
                dst[ia + 1] = f.s0b;
            }
        }
        pa = ma;
        pb = mb;
        ma = na;
        mb = nb;
    };
    for (int t = k0 - 2; t < k1; t++) step(t, fetch(t));
    // workgroup sum in a fixed order
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
    __shared__ double wsum[kCgsWaves];
    if (lane == 0) wsum[threadIdx.y] = sq;
    __syncthreads();
    if (lane == 0 && threadIdx.y == 0) {
        double acc = 0.0;
        for (int w = 0; w < kCgsWaves; w++) acc += wsum[w];
        partial[blockIdx.x] = acc;
    }
}

inline CgsTiles cgs_tiles(const LevelDesc &L) {
    // tile tx owns columns tx * kCgsOwnX - 1 .. tx * kCgsOwnX + 60
    return CgsTiles{(L.N[0] + 1 + kCgsOwnX - 1) / kCgsOwnX, (L.N[1] + kCgsOwnY - 1) / kCgsOwnY,
                    (L.N[2] + kCgsPlanes - 1) / kCgsPlanes};
}

__global__ void k_swap(double *__restrict__ a, double *__restrict__ b, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
        const double x = a[i], y = b[i];
        a[i] = y;
        b[i] = x;
    }
}

// dst = src with the fill applied (copy of the whole array + fill in one launch)
__global__ void k_copy_bc(LevelDesc L, BCView V, int main_z, int kz, double *__restrict__ dst,
                          const double *__restrict__ src) {
    if (static_cast<int>(blockIdx.z) >= main_z) {
        auto val = [&](long long c) { return src[c]; };
        dev::bc_item<double, decltype(val), true>(L, V, item_index(main_z), dst, val);
        return;
    }
    int bx, by, bz;
    main_block(main_z, bx, by, bz);
    const int o = V.shell_filled ? 1 : 0;  // the inner box when the fill writes every ghost
    const int i = o + bx * BX + static_cast<int>(threadIdx.x);
    const int j = o + by * BY + static_cast<int>(threadIdx.y);
    if (i >= L.N[0] - o || j >= L.N[1] - o) return;
#pragma unroll
    for (int q = 0; q < kz; q++) {
        const int k = o + bz * kz + q;
        if (k >= L.N[2] - o) break;
        const long long idx = lin(L, i, j, k);
        const bool fluid =
            i >= 1 && j >= 1 && k >= 1 && i <= L.n[0] && j <= L.n[1] && k <= L.n[2] && !in_obstacle(L, i, j, k);
        if (!fluid && dev::written(V, idx)) continue;  // fluid cells are never written by the fill
        const double v = src[idx];
        dst[idx] = v;
        if (fluid) dev::face_ghosts(L, V, dst, idx, i, j, k, v);
    }
}

// One Jacobi sweep of the multigrid smoother and the fill after it
// (JacobiDiffuse.cpp:292-300 + VCycleMG.cpp:713).  first = 1: the smoother's
// initial copy is folded in, every cell of dst is written (inner cells with
// the update of src, filled cells by the fill, all others with src);
// first = 0: cells that are neither inner nor filled keep dst's own values.
// ZERO (with FIRST): src is zero everywhere and is not read, the update is
// w beta dsign b, the value the sweep gives for src = 0.
template <typename T, typename TB, bool FIRST, bool ZERO>
__global__ void k_jacobi_bc(LevelDesc L, BCView V, int main_z, int kz, T ax, T ay, T az, T beta, T dsign, T w,
                            T *__restrict__ dst, const T *__restrict__ src, const TB *__restrict__ b) {
    static_assert(FIRST || !ZERO, "a sweep from zero writes every cell");
    if (static_cast<int>(blockIdx.z) >= main_z) {
        auto val = [&](long long c) -> T {
            int ci, cj, ck;
            if (is_inner(L, c, ci, cj, ck)) {
                if (ZERO) return w * (beta * (dsign * static_cast<T>(b[c])));
                return dev::jacobi_t<T, TB>(L, ax, ay, az, beta, dsign, w, src, b, c);
            }
            if (ZERO) return T(0);
            return FIRST ? src[c] : dst[c];
        };
        dev::bc_item<T, decltype(val), true>(L, V, item_index(main_z), dst, val);
        return;
    }
    int bx, by, bz;
    main_block(main_z, bx, by, bz);
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2], N0 = L.N[0], N1 = L.N[1];
    // a first sweep covers the whole array unless the fill writes every
    // ghost cell (then the inner box holds every cell it must write)
    const bool full = FIRST && !V.shell_filled;
    const int o = full ? 0 : 1;
    const int i = o + bx * BX + static_cast<int>(threadIdx.x);
    const int j = o + by * BY + static_cast<int>(threadIdx.y);
    if (full ? (i >= N0 || j >= N1) : (i > n0 || j > n1)) return;
    const int kend = full ? L.N[2] : n2 + 1;
    // 32-bit offsets (a level has fewer than 2^31 cells)
    const unsigned nj = static_cast<unsigned>(N0), nk = nj * static_cast<unsigned>(N1);
#pragma unroll
    for (int q = 0; q < kz; q++) {
        const int k = o + bz * kz + q;
        if (k >= kend) break;
        const unsigned idx = static_cast<unsigned>(i) + nj * static_cast<unsigned>(j) + nk * static_cast<unsigned>(k);
        const bool in_inner_box = full ? (i >= 1 && j >= 1 && k >= 1 && i <= n0 && j <= n1 && k <= n2) : true;
        // the update is computed for every cell of the inner box before the
        // obstacle test, so the stencil loads are in flight while it runs
        // (the expression of JacobiDiffuse.cpp:292-300, as dev::jacobi_t)
        T v = T(0);
        if (in_inner_box) {
            if (ZERO) {
                v = w * (beta * (dsign * static_cast<T>(b[idx])));
            } else {
                const T out_h = beta * (dsign * static_cast<T>(b[idx]) + ax * (src[idx + 1] + src[idx - 1]) +
                                        ay * (src[idx + nj] + src[idx - nj]) + az * (src[idx + nk] + src[idx - nk]));
                v = (1 - w) * src[idx] + w * out_h;
            }
        }
        if (in_inner_box && !in_obstacle(L, i, j, k)) {
            dst[idx] = v;
            dev::face_ghosts(L, V, dst, static_cast<long long>(idx), i, j, k, v);
        } else if (FIRST && !dev::written(V, idx)) {
            dst[idx] = ZERO ? T(0) : src[idx];
        }
    }
}

// Residuum (VCycleMG.cpp:346-354) and the obstacle-face part of the fill
// after it (VCycleMG.cpp:245); the ghosts of the residuum are never read.
template <typename T, typename TB>
__global__ void k_residuum_faces(LevelDesc L, BCView V, int main_z, int kz, T *__restrict__ out, const T *__restrict__ x,
                                 const TB *__restrict__ b) {
    if (static_cast<int>(blockIdx.z) >= main_z) {
        const long long t = item_index(main_z);
        if (t >= V.n_faces) return;
        dev::bc_item(L, V, t, out, [&](long long c) -> T {
            int ci, cj, ck;
            if (is_inner(L, c, ci, cj, ck)) return static_cast<T>(b[c]) - dev::laplace_t<T>(L, x, c);
            return out[c];
        });
        return;
    }
    int bx, by, bz;
    main_block(main_z, bx, by, bz);
    const int i = 1 + bx * BX + static_cast<int>(threadIdx.x);
    const int j = 1 + by * BY + static_cast<int>(threadIdx.y);
    if (i > L.n[0] || j > L.n[1]) return;
#pragma unroll
    for (int q = 0; q < kz; q++) {
        const int k = 1 + bz * kz + q;
        if (k > L.n[2]) break;
        if (in_obstacle(L, i, j, k)) continue;
        const long long idx = lin(L, i, j, k);
        out[idx] = static_cast<T>(b[idx]) - dev::laplace_t<T>(L, x, idx);
    }
}

// The correction of the ascent (VCycleMG.cpp:266-270) and the two fills at
// the start of the next smoothing, written out of place: dst = fine +
// prolongate(coarse) on the fine cells, with the fill applied.  Cells the
// prolongation does not reach get fine + 0.0, as the reference's whole-array
// addition of a zero correction gives.
template <typename Tf, typename Tc>
__device__ inline Tf corrected_at(const LevelDesc &Lc, const LevelDesc &Lf, const Tf *__restrict__ fine,
                                  const Tc *__restrict__ coarse, int i, int j, int k, long long c) {
    Tf p = 0.0;
    if (i >= 1 && j >= 1 && k >= 1) {
        const int ci = (i + 1) / 2, cj = (j + 1) / 2, ck = (k + 1) / 2;
        if (ci <= Lc.n[0] && cj <= Lc.n[1] && ck <= Lc.n[2] && !in_obstacle(Lc, ci, cj, ck)) {
            const int q = ((i & 1) ? 4 : 0) | ((j & 1) ? 2 : 0) | ((k & 1) ? 1 : 0);
            p = static_cast<Tf>(dev::prolong_t<Tc>(Lc, coarse, lin(Lc, ci, cj, ck), q));
        }
    }
    return fine[c] + p;
}
template <typename Tf, typename Tc>
__device__ inline Tf corrected(const LevelDesc &Lc, const LevelDesc &Lf, const Tf *__restrict__ fine,
                               const Tc *__restrict__ coarse, long long c) {
    const unsigned u = static_cast<unsigned>(c), Nx = static_cast<unsigned>(Lf.N[0]),
                   Ny = static_cast<unsigned>(Lf.N[1]);
    return corrected_at<Tf, Tc>(Lc, Lf, fine, coarse, static_cast<int>(u % Nx), static_cast<int>((u / Nx) % Ny),
                                static_cast<int>(u / (Nx * Ny)), c);
}
template <typename Tf, typename Tc>
__global__ void k_prolongate_add_bc(LevelDesc Lc, LevelDesc Lf, BCView V, int main_z, int kz, Tf *__restrict__ dst,
                                    const Tf *__restrict__ fine, const Tc *__restrict__ coarse) {
    if (static_cast<int>(blockIdx.z) >= main_z) {
        auto val = [&](long long c) { return corrected<Tf, Tc>(Lc, Lf, fine, coarse, c); };
        dev::bc_item<Tf, decltype(val), true>(Lf, V, item_index(main_z), dst, val);
        return;
    }
    int bx, by, bz;
    main_block(main_z, bx, by, bz);
    const int o = V.shell_filled ? 1 : 0;  // the inner box when the fill writes every ghost
    const int i = o + bx * BX + static_cast<int>(threadIdx.x);
    const int j = o + by * BY + static_cast<int>(threadIdx.y);
    if (i >= Lf.N[0] - o || j >= Lf.N[1] - o) return;
#pragma unroll
    for (int qz = 0; qz < kz; qz++) {
        const int k = o + bz * kz + qz;
        if (k >= Lf.N[2] - o) break;
        // 32-bit offsets (a level has fewer than 2^31 cells); the coarse and
        // fine loads are issued before the obstacle and fill tests
        const unsigned idx = static_cast<unsigned>(i) +
                             static_cast<unsigned>(Lf.N[0]) * (static_cast<unsigned>(j) +
                                                               static_cast<unsigned>(Lf.N[1]) * static_cast<unsigned>(k));
        const int ci = (i + 1) / 2, cj = (j + 1) / 2, ck = (k + 1) / 2;
        const bool reach = i >= 1 && j >= 1 && k >= 1 && ci <= Lc.n[0] && cj <= Lc.n[1] && ck <= Lc.n[2];
        const Tf f = fine[idx];
        Tf p = 0.0;
        if (reach) {
            const int q = ((i & 1) ? 4 : 0) | ((j & 1) ? 2 : 0) | ((k & 1) ? 1 : 0);
            const unsigned cidx = static_cast<unsigned>(ci) +
                                  static_cast<unsigned>(Lc.N[0]) *
                                      (static_cast<unsigned>(cj) + static_cast<unsigned>(Lc.N[1]) * static_cast<unsigned>(ck));
            const Tf pc = static_cast<Tf>(dev::prolong_t<Tc>(Lc, coarse, cidx, q));
            if (!in_obstacle(Lc, ci, cj, ck)) p = pc;
        }
        // fluid cells are never written by the fill: the write-bit lookup
        // is only needed for the others
        const bool fluid =
            i >= 1 && j >= 1 && k >= 1 && i <= Lf.n[0] && j <= Lf.n[1] && k <= Lf.n[2] && !in_obstacle(Lf, i, j, k);
        if (!fluid && dev::written(V, idx)) continue;
        const Tf v = f + p;  // corrected_at's value
        dst[idx] = v;
        if (fluid) dev::face_ghosts(Lf, V, dst, static_cast<long long>(idx), i, j, k, v);
    }
}

// x[cells[r]] = sum_j M[r][j] b[cells[j]]: one wavefront per row, the
// gathered right-hand side staged in LDS, a fixed summation order
__global__ void k_dense_apply(int n, const double *__restrict__ M, const uint32_t *__restrict__ cells,
                              const double *__restrict__ b, double *__restrict__ x) {
    extern __shared__ double sb[];
    for (int j = static_cast<int>(threadIdx.x); j < n; j += static_cast<int>(blockDim.x)) sb[j] = b[cells[j]];
    __syncthreads();
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int row = static_cast<int>(blockIdx.x) * (static_cast<int>(blockDim.x) >> 6) +
                    (static_cast<int>(threadIdx.x) >> 6);
    if (row >= n) return;
    const double *m = M + static_cast<size_t>(row) * static_cast<size_t>(n);
    double acc = 0.0;
    for (int j = lane; j < n; j += 64) acc += m[j] * sb[j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) x[cells[row]] = acc;
}

// Restriction of the fine residuum to a coarse level fused with that level's
// first smoothing sweep from zero (k_restrict_fill, then k_jacobi_bc with
// ZERO): rhs = restrict(fine) on the coarse fluid cells, err = w beta dsign
// rhs there, err's fill, zero on every other cell.  Fill items restrict the
// cell they read themselves, so no cell waits for another workgroup.
__global__ void k_restrict_sweep0(LevelDesc Lc, LevelDesc Lf, BCView V, int main_z, float w, float beta, float dsign,
                                  float *__restrict__ rhs, float *__restrict__ err, const float *__restrict__ fine) {
    if (static_cast<int>(blockIdx.z) >= main_z) {
        auto val = [&](long long c) -> float {
            int ci, cj, ck;
            if (!is_inner(Lc, c, ci, cj, ck)) return 0.0f;
            return w * (beta * (dsign * dev::restrict_t<float>(Lf, fine, ci, cj, ck)));
        };
        dev::bc_item<float, decltype(val), true>(Lc, V, item_index(main_z), err, val);
        return;
    }
    int bx, by, bz;
    main_block(main_z, bx, by, bz);
    const int i = bx * BX + static_cast<int>(threadIdx.x);
    const int j = by * BY + static_cast<int>(threadIdx.y);
    const int k = bz;
    if (i >= Lc.N[0] || j >= Lc.N[1]) return;
    const long long idx = lin(Lc, i, j, k);
    if (i >= 1 && j >= 1 && k >= 1 && i <= Lc.n[0] && j <= Lc.n[1] && k <= Lc.n[2] && !in_obstacle(Lc, i, j, k)) {
        const float r = dev::restrict_t<float>(Lf, fine, i, j, k);
        rhs[idx] = r;
        const float v = w * (beta * (dsign * r));
        err[idx] = v;
        dev::face_ghosts(Lc, V, err, idx, i, j, k, v);
    } else if (!dev::written(V, idx)) {
        err[idx] = 0.0f;
    }
}

// One idempotent fill in a single launch (faces, ghosts, dependent ghosts)
template <typename T>
__global__ void k_fill_bc(LevelDesc L, BCView V, T *__restrict__ f) {
    dev::bc_item(L, V, blockIdx.x * 256LL + threadIdx.x, f, [&](long long c) { return f[c]; });
}

// the fixed-order sum of k_reduce_sum, then the loop test of the diffusion
// iteration (ColoredGaussSeidelDiffuse.cpp:60-98): run again while
// sqrt(sum) > tol and fewer than max_iter iterations ran
__global__ void k_reduce_gate(const double *__restrict__ partial, int n, IterGate *__restrict__ g) {
    if (g->done) return;
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partial[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __shared__ double wsum[16];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; w++) r += wsum[w];
        const int c = g->count + 1;
        g->sum = r;
        g->count = c;
        g->done = (sqrt(r) > g->tol && c < g->max_iter) ? 0 : 1;
    }
}
__global__ void k_log_gate(const IterGate *__restrict__ g, LoopLog *__restrict__ log) {
    if (threadIdx.x == 0) {
        log->value = sqrt(g->sum);
        log->count = g->count;
        log->ok = g->done;
        log->start = 0.0;
    }
}
__global__ void k_alternate_finish(const IterGate *__restrict__ g, double *__restrict__ out,
                                   double *__restrict__ in, long long n) {
    const int c = g->count;
    if (c > 0 && (c & 1)) return;  // uniform over the grid
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
        const double x = out[i], y = in[i];
        out[i] = y;
        if (c != 0) in[i] = x;
    }
}
__global__ void k_gate_open_iter(IterGate *__restrict__ g, double tol, int max_iter) {
    if (threadIdx.x == 0) {
        g->sum = 1.0;
        g->tol = tol;
        g->count = 0;
        g->max_iter = max_iter;
        g->done = (1.0 > tol && 0 < max_iter) ? 0 : 1;
    }
}

__global__ void k_reduce_sum(const double *__restrict__ partial, int n, double *__restrict__ out) {
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partial[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __shared__ double wsum[16];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; w++) r += wsum[w];
        out[0] = r;
    }
}

// ------------------------------------------------------------------ IPressure
// IPressure.cpp:58-104: divergence on inner cells, zero on the ghost shell
__global__ void k_divergence(LevelDesc L, double hx, double hy, double hz, double *__restrict__ out,
                             const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z) {
    int bx, by, bz;
    main_block(static_cast<int>(gridDim.z), bx, by, bz);
    const int i = bx * BX + static_cast<int>(threadIdx.x);
    const int j = by * BY + static_cast<int>(threadIdx.y);
    const int k = bz;
    if (i >= L.N[0] || j >= L.N[1]) return;
    const long long idx = lin(L, i, j, k);
    if (i == 0 || j == 0 || k == 0 || i == L.N[0] - 1 || j == L.N[1] - 1 || k == L.N[2] - 1) {
        out[idx] = 0;
        return;
    }
    if (in_obstacle(L, i, j, k)) return;
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const double vx = hx * (x[idx + 1] - x[idx - 1]);
    const double vy = hy * (y[idx + nj] - y[idx - nj]);
    const double vz = hz * (z[idx + nk] - z[idx - nk]);
    out[idx] = vx + vy + vz;
}

// IPressure.cpp:154-161
__global__ void k_projection(LevelDesc L, double hx, double hy, double hz, double *__restrict__ ou,
                             double *__restrict__ ov, double *__restrict__ ow, const double *__restrict__ iu,
                             const double *__restrict__ iv, const double *__restrict__ iw,
                             const double *__restrict__ p) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    ou[idx] = iu[idx] - hx * (p[idx + 1] - p[idx - 1]);
    ov[idx] = iv[idx] - hy * (p[idx + nj] - p[idx - nj]);
    ow[idx] = iw[idx] - hz * (p[idx + nk] - p[idx - nk]);
}

// ------------------------------------------------------------------ ConstSmagorinsky
__global__ void k_smagorinsky(LevelDesc L, double rdx, double rdy, double rdz, double cdd, double *__restrict__ ev,
                              const double *__restrict__ u, const double *__restrict__ v,
                              const double *__restrict__ w) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long ni = 1, nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const double S11 = (u[idx + ni] - u[idx - ni]) * 0.5 * rdx;
    const double S22 = (v[idx + nj] - v[idx - nj]) * 0.5 * rdy;
    const double S33 = (w[idx + nk] - w[idx - nk]) * 0.5 * rdz;
    const double S12 = 0.5 * ((u[idx + nj] - u[idx - nj]) * 0.5 * rdy + (v[idx + ni] - v[idx - ni]) * 0.5 * rdx);
    const double S13 = 0.5 * ((u[idx + nk] - u[idx - nk]) * 0.5 * rdz + (w[idx + ni] - w[idx - ni]) * 0.5 * rdx);
    const double S23 = 0.5 * ((v[idx + nk] - v[idx - nk]) * 0.5 * rdz + (w[idx + nj] - w[idx - nj]) * 0.5 * rdy);
    const double S_bar = sqrt(2. * (S11 * S11 + S22 * S22 + S33 * S33 + 2. * (S12 * S12) + 2. * (S13 * S13) +
                                    2. * (S23 * S23)));
    ev[idx] = cdd * S_bar;
}

// ------------------------------------------------------------------ sources
__global__ void k_add_source(LevelDesc L, double dt, double *__restrict__ out, const double *__restrict__ s) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    out[idx] += dt * s[idx];
}

__global__ void k_buoyancy(LevelDesc L, double beta, double g, double *__restrict__ out,
                           const double *__restrict__ T, const double *__restrict__ Ta) {
    const int i = static_cast<int>(blockIdx.x) * BX + static_cast<int>(threadIdx.x);
    const int j = static_cast<int>(blockIdx.y) * BY + static_cast<int>(threadIdx.y);
    const int k = static_cast<int>(blockIdx.z);
    if (i >= L.N[0] || j >= L.N[1]) return;
    const bool ghost = i == 0 || j == 0 || k == 0 || i == L.N[0] - 1 || j == L.N[1] - 1 || k == L.N[2] - 1;
    if (!ghost && in_obstacle(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    out[idx] = -beta * (T[idx] - Ta[idx]) * g;
}

}  // namespace

// ================================================================== launchers
void copy(hipStream_t s, double *dst, const double *src, long long n, const int *gate) {
    k_copy<<<grid1d(n, 256), 256, 0, s>>>(dst, src, n, gate);
}
void dense_apply(hipStream_t s, int n, const double *M, const uint32_t *cells, const double *b, double *x) {
    if (n < 1 || n > kMaxDenseRows) throw std::invalid_argument("dense_apply: 1 to kMaxDenseRows rows");
    k_dense_apply<<<(n + 3) / 4, 256, sizeof(double) * static_cast<size_t>(n), s>>>(n, M, cells, b, x);
}
void copy_fan(hipStream_t s, const CopyFan &f, long long n) {
    switch (f.fields) {
    case 1: k_copy_fan<1><<<grid1d(n, 256), 256, 0, s>>>(f, n); break;
    case 2: k_copy_fan<2><<<grid1d(n, 256), 256, 0, s>>>(f, n); break;
    case 3: k_copy_fan<3><<<grid1d(n, 256), 256, 0, s>>>(f, n); break;
    case 4: k_copy_fan<4><<<grid1d(n, 256), 256, 0, s>>>(f, n); break;
    default: throw std::invalid_argument("copy_fan: 1 to 4 fields");
    }
}
void fill(hipStream_t s, double *dst, double v, long long n) { k_fill<<<grid1d(n, 256), 256, 0, s>>>(dst, v, n); }
void fill32(hipStream_t s, float *dst, float v, long long n) { k_fill32<<<grid1d(n, 256), 256, 0, s>>>(dst, v, n); }
void scale(hipStream_t s, double *dst, double v, long long n) { k_scale<<<grid1d(n, 256), 256, 0, s>>>(dst, v, n); }
void add(hipStream_t s, double *dst, const double *src, long long n) {
    k_add<<<grid1d(n, 256), 256, 0, s>>>(dst, src, n);
}
void scale_to(hipStream_t s, double *dst, const double *src, double v, long long n) {
    k_scale_to<<<grid1d(n, 256), 256, 0, s>>>(dst, src, v, n);
}
void copy_many(hipStream_t s, double *const *dst, const double *const *src, int count, long long n) {
    if (count <= 0) return;
    const unsigned g = std::min(grid1d(n, 256), 2048u);
    k_copy_many<<<dim3(g, static_cast<unsigned>(count)), 256, 0, s>>>(dst, src, n);
}
void scale_to_dev(hipStream_t s, double *dst, const double *src, const double *v, long long n) {
    k_scale_to_dev<<<grid1d(n, 256), 256, 0, s>>>(dst, src, v, n);
}
void copy_domain(hipStream_t s, const LevelDesc &L, double *dst, const double *src) {
    k_copy_domain<<<full_grid(L), block(), 0, s>>>(L, dst, src);
}

void bc_faces(hipStream_t s, const FaceEntry *e, long long n, const FaceParam *p, double *f, const int *gate) {
    if (n <= 0) return;
    k_bc_faces<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(e, n, p, f, gate);
}
void bc_domain(hipStream_t s, const LevelDesc &L, const DomainBC &bc, double *f, const int *gate) {
    if (!bc.active) return;
    const long long n = dev::ghost_count(L);
    k_bc_domain<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(L, bc, f, gate);
}

void bc_merged(hipStream_t s, const LevelDesc &L, const DomainBC &bc, const FaceEntry *e, long long n_faces,
               const FaceParam *p, double *f) {
    const long long n = n_faces + (bc.active ? dev::ghost_count(L) : 0);
    if (n <= 0) return;
    k_bc_merged<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(L, bc, e, n_faces, p, f);
}

void sl_advect(hipStream_t s, const LevelDesc &L, double dt, double *out, const double *in, const double *u,
               const double *v, const double *w) {
    k_sl_advect<<<inner_grid(L), block(), 0, s>>>(L, dt / L.h[0], dt / L.h[1], dt / L.h[2], out, in, u, v, w);
}

void jacobi_diffuse_step(hipStream_t s, const LevelDesc &L, double *out, const double *in, const double *b,
                         double ax, double ay, double az, double rb, double D, const double *ev, double dt,
                         double w, double *partial) {
    DiffCoef c{ax, ay, az, rb, D, dt, 1. / L.h[0], 1. / L.h[1], 1. / L.h[2]};
    k_jacobi_diffuse<<<inner_grid(L), block(), 0, s>>>(L, c, w, out, in, b, ev, partial);
}
void jacobi_step_mg(hipStream_t s, const LevelDesc &L, double *out, const double *in, const double *b, double ax,
                    double ay, double az, double beta, double dsign, double w) {
    k_jacobi_mg<<<inner_grid(L), block(), 0, s>>>(L, ax, ay, az, beta, dsign, w, out, in, b);
}

void jacobi_mg_first(hipStream_t s, const LevelDesc &L, double *dst, const double *src, const double *b, double ax,
                     double ay, double az, double beta, double dsign, double w) {
    k_jacobi_mg_first<<<full_grid(L), block(), 0, s>>>(L, ax, ay, az, beta, dsign, w, dst, src, b);
}

void cgs_color(hipStream_t s, const LevelDesc &L, int color, double *out, const double *b, double ax, double ay,
               double az, double beta, double dsign, double w, double D, const double *ev, double dt,
               const int *gate) {
    const int half = (L.n[0] + 1) / 2;
    dim3 g((half + BX - 1) / BX, (L.n[1] + BY - 1) / BY, L.n[2]);
    k_cgs_color<<<g, block(), 0, s>>>(L, color, ax, ay, az, beta, dsign, w, D, ev, dt, 1. / L.h[0], 1. / L.h[1],
                                      1. / L.h[2], out, b, gate);
}
void update_norm_partials(hipStream_t s, const LevelDesc &L, const double *out, const double *in, double rb,
                          double D, const double *ev, double dt, double *partial, const int *gate) {
    DiffCoef c{0, 0, 0, rb, D, dt, 1. / L.h[0], 1. / L.h[1], 1. / L.h[2]};
    k_update_norm<<<inner_grid(L), block(), 0, s>>>(L, c, out, in, ev, partial, gate);
}

int cgs_iter_partials(const LevelDesc &L) {
    // k_cgs_iter reads each lane's two cells with one 16-byte load, which
    // must not straddle the end of a row: an even row length; its byte
    // offsets, up to two planes past the array, must fit 31 bits
    const long long plane = static_cast<long long>(L.N[0]) * L.N[1];
    if (L.N[0] % 2 != 0 || (L.size + 3 * plane) * 8 >= (1LL << 31)) return std::numeric_limits<int>::max();
    const CgsTiles t = cgs_tiles(L);
    return t.tx * t.ty * t.tz;
}
void cgs_iter(hipStream_t s, const LevelDesc &L, const BCView &V, const CgsCoef &c, const double *ev, const double *b,
              const double *src, double *dst, double *partial, const int *gate) {
    const CgsTiles t = cgs_tiles(L);
    const long long main_blocks = static_cast<long long>(t.tx) * t.ty * t.tz;
    const long long items = dev::bc_items_grid(V), per = 64 * kCgsWaves;
    const unsigned g = static_cast<unsigned>(main_blocks + (items + per - 1) / per);
    if (ev != nullptr)
        k_cgs_iter<true><<<g, dim3(64, kCgsWaves), 0, s>>>(L, V, t, c, ev, b, src, dst, partial, gate);
    else
        k_cgs_iter<false><<<g, dim3(64, kCgsWaves), 0, s>>>(L, V, t, c, ev, b, src, dst, partial, gate);
}
void swap_arrays(hipStream_t s, double *a, double *b, long long n) {
    k_swap<<<grid1d(n, 256), 256, 0, s>>>(a, b, n);
}
void residuum(hipStream_t s, const LevelDesc &L, double *out, const double *x, const double *b) {
    k_residuum<<<inner_grid(L), block(), 0, s>>>(L, out, x, b);
}
void residual_partials(hipStream_t s, const LevelDesc &L, const double *x, const double *b, double *partial) {
    k_residual_partials<<<inner_grid(L), block(), 0, s>>>(L, x, b, partial);
}
void restrict_(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, double *coarse, const double *fine) {
    k_restrict<<<inner_grid(Lc), block(), 0, s>>>(Lc, Lf, coarse, fine);
}
void prolongate(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, double *fine, const double *coarse) {
    k_prolongate<<<inner_grid(Lc), block(), 0, s>>>(Lc, Lf, fine, coarse);
}

template <typename Tin, typename Tout, typename Terr>
void restrict_fill(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, Tout *coarse, const Tin *fine,
                   Terr *coarse_err) {
    k_restrict_fill<Tin, Tout, Terr><<<full_grid(Lc), block(), 0, s>>>(Lc, Lf, coarse, fine, coarse_err);
}
template void restrict_fill<double, double, double>(hipStream_t, const LevelDesc &, const LevelDesc &, double *,
                                                    const double *, double *);
template void restrict_fill<float, float, float>(hipStream_t, const LevelDesc &, const LevelDesc &, float *,
                                                 const float *, float *);
template void restrict_fill<float, double, double>(hipStream_t, const LevelDesc &, const LevelDesc &, double *,
                                                   const float *, double *);
void restrict_sweep0(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, const BCView &V, float *rhs, float *err,
                     const float *fine, double ax, double ay, double az, double beta, double dsign, double w) {
    (void)ax;
    (void)ay;
    (void)az;
    const dim3 g = full_grid(Lc);
    k_restrict_sweep0<<<with_items(g, dev::bc_items_grid(V)), block(), 0, s>>>(
        Lc, Lf, V, static_cast<int>(g.z), static_cast<float>(w), static_cast<float>(beta), static_cast<float>(dsign), rhs,
        err, fine);
}
void prolongate_add(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, double *fine, const double *coarse) {
    k_prolongate_add<<<inner_grid(Lc), block(), 0, s>>>(Lc, Lf, fine, coarse);
}

// Times the level's FP32 first sweep with 1, 2 and 4 cells per thread along
// z on scratch arrays of the level's size and keeps the fastest for every
// fused level kernel of that size (the values do not depend on the choice).
// Not during stream capture: it synchronizes on events.
void tune_level(hipStream_t s, const LevelDesc &L, const BCView &V) {
    const TuneKey key = tune_key(L, V);
    {
        std::lock_guard<std::mutex> g(kz_lock);
        if (kz_tuned.count(key)) return;
    }
    const size_t n = static_cast<size_t>(L.size);
    float *buf = nullptr;
    ARTSS_HIP(hipMalloc(&buf, sizeof(float) * 3 * n));
    ARTSS_HIP(hipMemsetAsync(buf, 0, sizeof(float) * 3 * n, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ARTSS_HIP(hipEventCreate(&e0));
    ARTSS_HIP(hipEventCreate(&e1));
    const int cand[3] = {1, 2, KZ};
    int best = L.size >= kBatchCells ? KZ : 1;
    float best_ms = 0.0f;
    const dim3 g0 = V.shell_filled ? inner_grid(L) : full_grid(L);
    for (int c : cand) {
        const dim3 g = zbatch(g0, c);
        const dim3 gi = with_items(g, dev::bc_items_grid(V));
        auto launch = [&] {
            k_jacobi_bc<float, float, true, false><<<gi, block(), 0, s>>>(L, V, static_cast<int>(g.z), c, 1.f, 1.f, 1.f,
                                                                         0.16f, -1.f, 0.8f, buf, buf + n, buf + 2 * n);
        };
        launch();  // warm-up
        ARTSS_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 5; r++) launch();
        ARTSS_HIP(hipEventRecord(e1, s));
        ARTSS_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        ARTSS_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (c == cand[0] || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ARTSS_HIP(hipStreamSynchronize(s));
    ARTSS_HIP(hipFree(buf));
    std::lock_guard<std::mutex> g(kz_lock);
    kz_tuned.emplace(key, best);
}
int tuned_kz(long long level_size) {
    std::lock_guard<std::mutex> g(kz_lock);
    for (const auto &kv : kz_tuned)
        if (kv.first[0] * kv.first[1] * kv.first[2] == level_size) return kv.second;
    return 0;
}
void set_forced_kz(int kz) {
    if (kz != 0 && kz != 1 && kz != 2 && kz != KZ)
        throw std::invalid_argument("cells per thread along z must be 0 (autotuned), 1, 2 or 4");
    kz_forced.store(kz);
}

void copy_bc(hipStream_t s, const LevelDesc &L, const BCView &V, double *dst, const double *src) {
    const int kz = kz_of(L, V);
    const dim3 g = zbatch(V.shell_filled ? inner_grid(L) : full_grid(L), kz);
    k_copy_bc<<<with_items(g, dev::bc_items_grid(V)), block(), 0, s>>>(L, V, static_cast<int>(g.z), kz, dst, src);
}
template <typename T, typename TB>
void jacobi_bc(hipStream_t s, const LevelDesc &L, const BCView &V, bool first, T *dst, const T *src, const TB *b,
               double ax, double ay, double az, double beta, double dsign, double w) {
    if (src == nullptr && !first) throw std::invalid_argument("jacobi_bc: a sweep from zero must be a first sweep");
    const int kz = kz_of(L, V);
    const dim3 g = zbatch(first && !V.shell_filled ? full_grid(L) : inner_grid(L), kz);
    const dim3 gi = with_items(g, dev::bc_items_grid(V));
    const T ax_ = static_cast<T>(ax), ay_ = static_cast<T>(ay), az_ = static_cast<T>(az), beta_ = static_cast<T>(beta),
            ds_ = static_cast<T>(dsign), w_ = static_cast<T>(w);
    if (first && src == nullptr)
        k_jacobi_bc<T, TB, true, true><<<gi, block(), 0, s>>>(L, V, static_cast<int>(g.z), kz, ax_, ay_, az_, beta_, ds_,
                                                              w_, dst, src, b);
    else if (first)
        k_jacobi_bc<T, TB, true, false><<<gi, block(), 0, s>>>(L, V, static_cast<int>(g.z), kz, ax_, ay_, az_, beta_, ds_,
                                                               w_, dst, src, b);
    else
        k_jacobi_bc<T, TB, false, false><<<gi, block(), 0, s>>>(L, V, static_cast<int>(g.z), kz, ax_, ay_, az_, beta_,
                                                                ds_, w_, dst, src, b);
}
template void jacobi_bc<double, double>(hipStream_t, const LevelDesc &, const BCView &, bool, double *,
                                        const double *, const double *, double, double, double, double, double,
                                        double);
template void jacobi_bc<float, double>(hipStream_t, const LevelDesc &, const BCView &, bool, float *, const float *,
                                       const double *, double, double, double, double, double, double);
template void jacobi_bc<float, float>(hipStream_t, const LevelDesc &, const BCView &, bool, float *, const float *,
                                      const float *, double, double, double, double, double, double);

template <typename T, typename TB>
void residuum_faces(hipStream_t s, const LevelDesc &L, const BCView &V, T *out, const T *x, const TB *b) {
    const int kz = kz_of(L, V);
    const dim3 g = zbatch(inner_grid(L), kz);
    k_residuum_faces<T, TB><<<with_items(g, V.n_faces), block(), 0, s>>>(L, V, static_cast<int>(g.z), kz, out, x, b);
}
template void residuum_faces<double, double>(hipStream_t, const LevelDesc &, const BCView &, double *,
                                             const double *, const double *);
template void residuum_faces<float, double>(hipStream_t, const LevelDesc &, const BCView &, float *, const float *,
                                            const double *);
template void residuum_faces<float, float>(hipStream_t, const LevelDesc &, const BCView &, float *, const float *,
                                           const float *);

template <typename Tf, typename Tc>
void prolongate_add_bc(hipStream_t s, const LevelDesc &Lc, const LevelDesc &Lf, const BCView &V, Tf *dst,
                       const Tf *fine, const Tc *coarse) {
    const int kz = kz_of(Lf, V);
    const dim3 g = zbatch(V.shell_filled ? inner_grid(Lf) : full_grid(Lf), kz);
    k_prolongate_add_bc<Tf, Tc><<<with_items(g, dev::bc_items_grid(V)), block(), 0, s>>>(Lc, Lf, V, static_cast<int>(g.z), kz,
                                                                                     dst, fine, coarse);
}
template void prolongate_add_bc<double, double>(hipStream_t, const LevelDesc &, const LevelDesc &, const BCView &,
                                                double *, const double *, const double *);
template void prolongate_add_bc<float, double>(hipStream_t, const LevelDesc &, const LevelDesc &, const BCView &,
                                               float *, const float *, const double *);
template void prolongate_add_bc<float, float>(hipStream_t, const LevelDesc &, const LevelDesc &, const BCView &,
                                              float *, const float *, const float *);

template <typename T>
void fill_bc(hipStream_t s, const LevelDesc &L, const BCView &V, T *f) {
    const long long n = dev::bc_items(V);
    if (n <= 0) return;
    k_fill_bc<T><<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(L, V, f);
}
template void fill_bc<double>(hipStream_t, const LevelDesc &, const BCView &, double *);
template void fill_bc<float>(hipStream_t, const LevelDesc &, const BCView &, float *);

int partial_count(const LevelDesc &L) {
    const dim3 g = inner_grid(L);
    return static_cast<int>(g.x * g.y * g.z);
}
void reduce_gate(hipStream_t s, const double *partial, int n, IterGate *g) {
    k_reduce_gate<<<1, 1024, 0, s>>>(partial, n, g);
}
void gate_open(hipStream_t s, IterGate *g, double tol, int max_iter) { k_gate_open_iter<<<1, 64, 0, s>>>(g, tol, max_iter); }
void log_gate(hipStream_t s, const IterGate *g, LoopLog *log) { k_log_gate<<<1, 64, 0, s>>>(g, log); }
void alternate_finish(hipStream_t s, const IterGate *g, double *out, double *in, long long n) {
    k_alternate_finish<<<grid1d(n, 256), 256, 0, s>>>(g, out, in, n);
}
void reduce_sum(hipStream_t s, const double *partial, int n, double *out) {
    k_reduce_sum<<<1, 1024, 0, s>>>(partial, n, out);
}

void dissipate(hipStream_t s, const LevelDesc &L, double *out, const double *u, const double *v, const double *w,
               double nu, double dt) {
    k_dissipate<<<inner_grid(L), block(), 0, s>>>(L, 1. / L.h[0], 1. / L.h[1], 1. / L.h[2], nu, dt, out, u, v, w);
}
void explicit_diffuse(hipStream_t s, const LevelDesc &L, double *out, const double *in, double D, const double *ev,
                      double dt) {
    const double rdx = D / (L.h[0] * L.h[0]), rdy = D / (L.h[1] * L.h[1]), rdz = D / (L.h[2] * L.h[2]);
    k_explicit_diffuse<<<inner_grid(L), block(), 0, s>>>(L, rdx, rdy, rdz, D, dt, out, in, ev);
}
void divergence(hipStream_t s, const LevelDesc &L, double *out, const double *x, const double *y, const double *z) {
    const double hx = 0.5 * (1. / L.h[0]), hy = 0.5 * (1. / L.h[1]), hz = 0.5 * (1. / L.h[2]);
    k_divergence<<<full_grid(L), block(), 0, s>>>(L, hx, hy, hz, out, x, y, z);
}
void projection(hipStream_t s, const LevelDesc &L, double *ou, double *ov, double *ow, const double *iu,
                const double *iv, const double *iw, const double *p) {
    const double hx = 0.5 * (1. / L.h[0]), hy = 0.5 * (1. / L.h[1]), hz = 0.5 * (1. / L.h[2]);
    k_projection<<<inner_grid(L), block(), 0, s>>>(L, hx, hy, hz, ou, ov, ow, iu, iv, iw, p);
}

// ------------------------------------------------------------------ dynamic Smagorinsky
// DynamicSmagorinsky.cpp:172-427.  Every kernel writes the inner fluid cells
// of its outputs only; the other cells of the work fields stay zero.

// explicit_filtering (:405-427) of `count` arrays at once: the trapezoidal
// 3x3x3 filter summed with the x offset outermost, then y, then z
__global__ void k_dyn_filter(LevelDesc L, DynArrays out, DynConstArrays in, int count) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const double a[3] = {1. / 4., 1. / 2., 1. / 4.};
    for (int q = 0; q < count; q++) {
        const double *f = in.p[q];
        double sum = 0;
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
            for (int m = 0; m < 3; m++)
#pragma unroll
                for (int n = 0; n < 3; n++) sum = sum + a[l] * a[m] * a[n] * f[idx - 1 + l + nj * (m - 1) + nk * (n - 1)];
        out.p[q][idx] = sum;
    }
}

// :212-228: the six products of the filtered velocities (the reference forms
// both its "velocity products" and its "filtered velocity products" this way)
__global__ void k_dyn_products(LevelDesc L, const double *__restrict__ uf, const double *__restrict__ vf,
                               const double *__restrict__ wf, DynArrays prod) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const double u = uf[idx], v = vf[idx], w = wf[idx];
    prod.p[0][idx] = u * u;
    prod.p[1][idx] = v * v;
    prod.p[2][idx] = w * w;
    prod.p[3][idx] = u * v;
    prod.p[4][idx] = u * w;
    prod.p[5][idx] = v * w;
}

// :256-282: strain tensor of the filtered velocity, its modulus, and the
// products modulus x tensor (S11 S22 S33 S12 S13 S23 in out.p[0..5], the
// products in out.p[6..11])
__global__ void k_dyn_strain(LevelDesc L, double rdx, double rdy, double rdz, const double *__restrict__ uf,
                             const double *__restrict__ vf, const double *__restrict__ wf, DynArrays out,
                             double *__restrict__ sbar) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    double S[6];
    S[0] = (uf[idx + 1] - uf[idx - 1]) * 0.5 * rdx;
    S[1] = (vf[idx + nj] - vf[idx - nj]) * 0.5 * rdy;
    S[2] = (wf[idx + nk] - wf[idx - nk]) * 0.5 * rdz;
    S[3] = 0.5 * ((uf[idx + nj] - uf[idx - nj]) * 0.5 * rdy + (vf[idx + 1] - vf[idx - 1]) * 0.5 * rdx);
    S[4] = 0.5 * ((uf[idx + nk] - uf[idx - nk]) * 0.5 * rdz + (wf[idx + 1] - wf[idx - 1]) * 0.5 * rdx);
    S[5] = 0.5 * ((vf[idx + nk] - vf[idx - nk]) * 0.5 * rdz + (wf[idx + nj] - wf[idx - nj]) * 0.5 * rdy);
    const double sb = sqrt(2. * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2. * (S[3] * S[3]) + 2. * (S[4] * S[4]) +
                                 2. * (S[5] * S[5])));
    sbar[idx] = sb;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        out.p[c][idx] = S[c];
        out.p[6 + c][idx] = sb * S[c];
    }
}

// :308-350: Leonard stress, the filtered strain modulus, M, and the raw
// coefficient num / den with the reference's products where sums belong.
// filt.p[0..5]: filtered products uu vv ww uv uw vw (uw: never filtered, zero);
// prod.p[0..5]: the products; sf.p[0..5]: filtered strain, sf.p[6..11]:
// filtered modulus x strain.  c2 = 2 delta_s delta_s, a2 = alpha alpha.
__global__ void k_dyn_coef(LevelDesc L, double c2, double a2, DynConstArrays filt, DynConstArrays prod,
                           DynConstArrays sf, double *__restrict__ cs) {
    int i, j, k;
    if (!inner_cell(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    double Lt[6], M[6], S[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        Lt[c] = filt.p[c][idx] - prod.p[c][idx];
        S[c] = sf.p[c][idx];
    }
    const double sb = sqrt(2. * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2. * (S[3] * S[3]) + 2. * (S[4] * S[4]) +
                                 2. * (S[5] * S[5])));
#pragma unroll
    for (int c = 0; c < 6; c++) M[c] = c2 * (sf.p[6 + c][idx] - a2 * sb * S[c]);
    const double num = Lt[0] * M[0] + Lt[1] * M[1] * Lt[2] * M[2] + 2.0 * Lt[3] * M[3] + 2.0 * Lt[4] * M[4] +
                       2.0 * Lt[5] * M[5];
    const double den = M[0] * M[0] + M[1] * M[1] * M[2] * M[2] + 2.0 * M[3] * M[3] + 2.0 * M[4] * M[4] +
                       2.0 * M[5] * M[5];
    cs[idx] = num / den;
}

// :372-391 for the cells i + 2 j + 4 k = t: the in-place average over the
// 27 neighbours reads averaged values from every neighbour earlier in index
// order (all of them lie on smaller t) and raw values from the later ones
// (larger t), exactly as the sequential loop in index order does; then the
// eddy viscosity of the cell and its clamp.  One thread per (j, k).
__global__ void k_dyn_average(LevelDesc L, int t, double delta_s, double nu, double *__restrict__ cs,
                              const double *__restrict__ sbar, double *__restrict__ ev) {
    const int q = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (q >= L.n[1] * L.n[2]) return;
    const int j = 1 + q % L.n[1], k = 1 + q / L.n[1];
    const int i = t - 2 * j - 4 * k;
    if (i < 1 || i > L.n[0] || in_obstacle(L, i, j, k)) return;
    const long long idx = lin(L, i, j, k);
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    double sum = 0;
    for (int l = 0; l < 3; l++)
        for (int m = 0; m < 3; m++)
            for (int n = 0; n < 3; n++) sum = sum + cs[idx - 1 + l + nj * (m - 1) + nk * (n - 1)];
    const double c = sum / 27.0;
    cs[idx] = c;
    double e = 2.0 * c * delta_s * delta_s * sbar[idx];
    if ((e + nu) < 0) e = 0;
    ev[idx] = e;
}

void dyn_filter(hipStream_t s, const LevelDesc &L, const DynArrays &out, const DynConstArrays &in, int count) {
    if (count < 1 || count > kDynMax) throw std::invalid_argument("dyn_filter: 1 to 12 arrays");
    k_dyn_filter<<<inner_grid(L), block(), 0, s>>>(L, out, in, count);
}
void dyn_products(hipStream_t s, const LevelDesc &L, const double *uf, const double *vf, const double *wf,
                  const DynArrays &prod) {
    k_dyn_products<<<inner_grid(L), block(), 0, s>>>(L, uf, vf, wf, prod);
}
void dyn_strain(hipStream_t s, const LevelDesc &L, const double *uf, const double *vf, const double *wf,
                const DynArrays &out, double *sbar) {
    k_dyn_strain<<<inner_grid(L), block(), 0, s>>>(L, 1. / L.h[0], 1. / L.h[1], 1. / L.h[2], uf, vf, wf, out, sbar);
}
void dyn_coef(hipStream_t s, const LevelDesc &L, double c2, double a2, const DynConstArrays &filt,
              const DynConstArrays &prod, const DynConstArrays &sf, double *cs) {
    k_dyn_coef<<<inner_grid(L), block(), 0, s>>>(L, c2, a2, filt, prod, sf, cs);
}
void dyn_average(hipStream_t s, const LevelDesc &L, double delta_s, double nu, double *cs, const double *sbar,
                 double *ev) {
    const long long cols = static_cast<long long>(L.n[1]) * L.n[2];
    const unsigned g = static_cast<unsigned>((cols + 255) / 256);
    const int t_last = L.n[0] + 2 * L.n[1] + 4 * L.n[2];
    for (int t = 1 + 2 + 4; t <= t_last; t++) k_dyn_average<<<g, 256, 0, s>>>(L, t, delta_s, nu, cs, sbar, ev);
}

void smagorinsky(hipStream_t s, const LevelDesc &L, double *ev, const double *u, const double *v, const double *w,
                 double cdd) {
    k_smagorinsky<<<inner_grid(L), block(), 0, s>>>(L, 1. / L.h[0], 1. / L.h[1], 1. / L.h[2], cdd, ev, u, v, w);
}

void add_source(hipStream_t s, const LevelDesc &L, double *out, const double *src, double dt) {
    k_add_source<<<inner_grid(L), block(), 0, s>>>(L, dt, out, src);
}
void buoyancy(hipStream_t s, const LevelDesc &L, double *out, const double *T, const double *Ta, double beta,
              double g) {
    k_buoyancy<<<full_grid(L), block(), 0, s>>>(L, beta, g, out, T, Ta);
}

}  // namespace k
}  // namespace artss
