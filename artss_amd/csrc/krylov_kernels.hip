// Kernels of the Krylov-accelerated pressure solve (krylov.cpp).
//
// GCR with one previous direction, right-preconditioned by a V-cycle: per
// iteration w = A z (A: the discrete Laplacian with the linear part of the
// boundary fill), four dot products, one coefficient kernel and one fused
// update.  FP64 arithmetic throughout; the preconditioned direction z and
// the kept direction zp may be stored in FP32.
// Dot products are per-workgroup partial sums added by one workgroup in a
// fixed order, so runs are deterministic.
#include <cmath>

#include "krylov_kernels.h"
#include "stencil.h"

namespace artss {
namespace kk {

namespace {

constexpr int BX = 64;
constexpr int BY = 4;

inline dim3 inner_grid(const LevelDesc &L) {
    return dim3((L.n[0] + BX - 1) / BX, (L.n[1] + BY - 1) / BY, L.n[2]);
}
inline dim3 full_grid(const LevelDesc &L) {
    return dim3((L.N[0] + BX - 1) / BX, (L.N[1] + BY - 1) / BY, L.N[2]);
}
// the update gives each thread kUpdateZ consecutive cells along z (waves of
// one cell each are bound by the workgroup launch rate at level 0)
constexpr int kUpdateZ = 4;
inline dim3 update_grid(const LevelDesc &L) {
    const dim3 g = full_grid(L);
    return dim3(g.x, g.y, (g.z + kUpdateZ - 1) / kUpdateZ);
}

__device__ inline int block_id() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }
__device__ inline int block_count() { return gridDim.x * gridDim.y * gridDim.z; }

// workgroup sums of Q values per thread, written to partial[q * blocks + block]
template <int Q>
__device__ inline void write_partials(double (&v)[Q], double *partial) {
    __shared__ double wsum[Q][BY];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (threadIdx.x == 0) wsum[q][threadIdx.y] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            double s = 0.0;
            for (int w = 0; w < BY; w++) s += wsum[q][w];
            partial[q * block_count() + block_id()] = s;
        }
    }
}

// The hardware deals workgroups to the 8 XCDs round robin by linear id; the
// bijection id -> (id % 8) * per + id / 8 on the first 8 * per ids gives each
// XCD a contiguous run of blocks, so stencil neighbours along z are mostly in
// that XCD's own L2.  Partial sums stay indexed by the linear id.
__device__ inline void xcd_block(int &bx, int &by, int &bz) {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned per = gx * gy * gridDim.z / 8u;
    const unsigned nid = id < per * 8u ? (id % 8u) * per + id / 8u : id;
    bx = static_cast<int>(nid % gx);
    by = static_cast<int>((nid / gx) % gy);
    bz = static_cast<int>(nid / (gx * gy));
}
__device__ inline bool inner_at(const LevelDesc &L, int &i, int &j, int &k) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    i = 1 + bx * BX + static_cast<int>(threadIdx.x);
    j = 1 + by * BY + static_cast<int>(threadIdx.y);
    k = 1 + bz;
    return i <= L.n[0] && j <= L.n[1] && !in_obstacle(L, i, j, k);
}

template <typename T>
__device__ inline double lap(const LevelDesc &L, const T *__restrict__ x, long long i) {
    const long long nj = L.N[0], nk = static_cast<long long>(L.N[0]) * L.N[1];
    const double rdx2 = 1. / (L.h[0] * L.h[0]), rdy2 = 1. / (L.h[1] * L.h[1]), rdz2 = 1. / (L.h[2] * L.h[2]);
    const double c = x[i];
    return rdx2 * (double(x[i - 1]) - 2 * c + double(x[i + 1])) + rdy2 * (double(x[i - nj]) - 2 * c + double(x[i + nj])) +
           rdz2 * (double(x[i - nk]) - 2 * c + double(x[i + nk]));
}

// r = b - lap(x) on inner cells, partial (r, r)
__global__ void k_residual_vec(LevelDesc L, double *__restrict__ r, const double *__restrict__ x,
                               const double *__restrict__ b, double *__restrict__ partial) {
    int i, j, k;
    double v[1] = {0.0};
    if (inner_at(L, i, j, k)) {
        const long long idx = dev::lin(L, i, j, k);
        const double x_ = b[idx] - dev::laplace(L, x, idx);
        r[idx] = x_;
        v[0] = x_ * x_;
    }
    write_partials<1>(v, partial);
}

// w = lap(z) and wp = lap(zp) on inner cells (neither is stored: the update
// recomputes both); partials (w,w), (w,wp), (r,w), (r,wp).  A thread sums
// kDotsZ planes before the four workgroup reductions, which cost more than
// the loads.
constexpr int kDotsZ = 4;
inline dim3 dots_grid(const LevelDesc &L) {
    return dim3((L.n[0] + BX - 1) / BX, (L.n[1] + BY - 1) / BY, (L.n[2] + kDotsZ - 1) / kDotsZ);
}
template <typename Tz>
__global__ void k_lap_dots(LevelDesc L, const Tz *__restrict__ z, const Tz *__restrict__ zp,
                           const double *__restrict__ r, double *__restrict__ partial) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int i = 1 + bx * BX + static_cast<int>(threadIdx.x);
    const int j = 1 + by * BY + static_cast<int>(threadIdx.y);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i <= L.n[0] && j <= L.n[1]) {
        for (int q = 0; q < kDotsZ; q++) {
            const int k = 1 + bz * kDotsZ + q;
            if (k > L.n[2]) break;
            if (in_obstacle(L, i, j, k)) continue;
            const long long idx = dev::lin(L, i, j, k);
            const double x = lap(L, z, idx);
            const double p = lap(L, zp, idx), rr = r[idx];
            v[0] += x * x;
            v[1] += x * p;
            v[2] += rr * x;
            v[3] += rr * p;
        }
    }
    write_partials<4>(v, partial);
}

// coef[0] = beta, coef[1] = 1/|w - beta wp|, coef[2] = alpha
__global__ void k_gcr_coef(const double *__restrict__ partial, int blocks, double *__restrict__ coef,
                           const GcrGate *__restrict__ gate) {
    if (gate != nullptr && gate->done) return;
    __shared__ double s[4][16];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < blocks; t += 1024)
        for (int q = 0; q < 4; q++) acc[q] += partial[q * blocks + t];
    for (int q = 0; q < 4; q++) {
        double x = acc[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) s[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S[4];
        for (int q = 0; q < 4; q++) {
            S[q] = 0.0;
            for (int w = 0; w < 16; w++) S[q] += s[q][w];
        }
        double beta = S[1];
        double nw2 = S[0] - beta * beta;
        if (!(nw2 > 1e-6 * S[0])) {  // w (almost) parallel to wp: restart without it
            beta = 0.0;
            nw2 = S[0];
        }
        const double inv = nw2 > 0.0 ? 1.0 / sqrt(nw2) : 0.0;
        coef[0] = beta;
        coef[1] = inv;
        coef[2] = (S[2] - beta * S[3]) * inv;
    }
}

// w = lap(z), wp = lap(zp) on inner cells, 0 elsewhere; z' = (z - beta zp)/|w'|,
// w' = (w - beta wp)/|w'|; x += alpha z'; r -= alpha w'; zp_out = z' over
// the whole array; partial (r, r).  wp is the Laplacian of the stored zp,
// never a stored copy of the previous w': the correction of r is then the
// Laplacian of the correction of x whatever precision zp is kept in, so
// the recursive residual follows the true one.  zp_out is a different
// buffer from zp (the stencil reads zp's neighbours).
template <typename Tz>
__global__ void k_gcr_update(LevelDesc L, const double *__restrict__ coef, double *__restrict__ x,
                             double *__restrict__ r, const Tz *__restrict__ z, const Tz *__restrict__ zp,
                             Tz *__restrict__ zp_out, double *__restrict__ partial, const GcrGate *__restrict__ gate) {
    if (gate != nullptr && gate->done) return;  // uniform over the grid
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int i = bx * BX + static_cast<int>(threadIdx.x);
    const int j = by * BY + static_cast<int>(threadIdx.y);
    double v[1] = {0.0};
    if (i < L.N[0] && j < L.N[1]) {
        const double beta = coef[0], inv = coef[1], alpha = coef[2];
#pragma unroll
        for (int q = 0; q < kUpdateZ; q++) {
            const int k = bz * kUpdateZ + q;
            if (k >= L.N[2]) break;
            const long long idx = dev::lin(L, i, j, k);
            const bool inner = i >= 1 && j >= 1 && k >= 1 && i <= L.n[0] && j <= L.n[1] && k <= L.n[2] &&
                               !in_obstacle(L, i, j, k);
            const double w = inner ? lap(L, z, idx) : 0.0;
            const double wp = inner ? lap(L, zp, idx) : 0.0;
            const double zn = (double(z[idx]) - beta * double(zp[idx])) * inv;
            const double wn = (w - beta * wp) * inv;
            x[idx] += alpha * zn;
            const double rn = r[idx] - alpha * wn;
            r[idx] = rn;
            zp_out[idx] = static_cast<Tz>(zn);
            v[0] += rn * rn;
        }
    }
    write_partials<1>(v, partial);
}

__global__ void k_sum(const double *__restrict__ partial, int n, double *__restrict__ out) {
    __shared__ double s[16];
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partial[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; w++) r += s[w];
        out[0] = r;
    }
}

// the sum of the update's partials closes the gate when (r, r) <= tol^2;
// nothing runs once it is closed, so the state stays that of the first
// iteration that met the tolerance
__global__ void k_sum_gate(const double *__restrict__ partial, int n, GcrGate *__restrict__ gate) {
    if (gate->done) return;
    __shared__ double s[16];
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partial[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; w++) r += s[w];
        gate->rr = r;
        gate->count += 1;
        gate->done = r <= gate->tol2 ? 1 : 0;
    }
}

// the start of a captured solve: (r, r) summed as k_sum does, the gate
// opened and closed at once when it already meets the tolerance
__global__ void k_gate_start(const double *__restrict__ partial, int n, GcrGate *__restrict__ gate, double tol2) {
    __shared__ double s[16];
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partial[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; w++) r += s[w];
        gate->rr = r;
        gate->rr0 = r;
        gate->tol2 = tol2;
        gate->count = 0;
        gate->done = r <= tol2 ? 1 : 0;
    }
}
// the end of a captured solve: the true residual's partials (the
// multigrid's residual_partials) summed as k::reduce_sum does; the solve
// ended as its eager form would when the gate closed and sqrt of that sum
// meets tol
__global__ void k_gate_confirm(const double *__restrict__ partial, int n, const GcrGate *__restrict__ gate,
                               double tol, LoopLog *__restrict__ log) {
    __shared__ double s[16];
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partial[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < 16; w++) r += s[w];
        const double t = sqrt(r);
        log->value = t;
        log->count = gate->count;
        log->ok = gate->done && t <= tol ? 1 : 0;
        log->start = sqrt(gate->rr0);
    }
}

__global__ void k_gate_open(GcrGate *__restrict__ gate, double tol2) {
    if (threadIdx.x == 0) {
        gate->rr = 0.0;
        gate->rr0 = 0.0;
        gate->tol2 = tol2;
        gate->count = 0;
        gate->done = 0;
    }
}

// ---- warm start
__global__ void k_warm_partials(LevelDesc L, const double *__restrict__ p, const double *__restrict__ p_prev,
                                double *__restrict__ partial) {
    int i, j, k;
    double v[1] = {0.0};
    if (inner_at(L, i, j, k)) {
        const long long idx = dev::lin(L, i, j, k);
        const double d = p[idx] - p_prev[idx];
        v[0] = d * d;
    }
    write_partials<1>(v, partial);
}
// 1 / rms of the increment over the fluid cells, 0 when there is none
__device__ inline double inv_rms(const double *sum, double n) { return *sum > 0.0 ? 1.0 / sqrt(*sum / n) : 0.0; }
__device__ inline long long box_index(const LevelDesc &L, int i, int j, int k) {
    return (i - 1) + static_cast<long long>(L.n[0]) * ((j - 1) + static_cast<long long>(L.n[1]) * (k - 1));
}
__global__ void k_warm_pack(LevelDesc L, const double *__restrict__ p, const double *__restrict__ b,
                            const double *__restrict__ p_prev, const double *__restrict__ sum, double n,
                            double c, float *__restrict__ x) {
    const int i = 1 + static_cast<int>(blockIdx.x) * BX + static_cast<int>(threadIdx.x);
    const int j = 1 + static_cast<int>(blockIdx.y) * BY + static_cast<int>(threadIdx.y);
    const int k = 1 + static_cast<int>(blockIdx.z);
    if (i > L.n[0] || j > L.n[1]) return;
    float r = 0.f, d = 0.f;
    if (!in_obstacle(L, i, j, k)) {
        const long long idx = dev::lin(L, i, j, k);
        const double q = inv_rms(sum, n);
        // residual of the extrapolated start 2 p - p_prev (the fill is affine,
        // so the filled arrays combine to the filled extrapolation)
        const double lap_e = 2.0 * dev::laplace(L, p, idx) - dev::laplace(L, p_prev, idx);
        r = static_cast<float>(c * (b[idx] - lap_e) * q);
        d = static_cast<float>((p[idx] - p_prev[idx]) * q);
    }
    const long long v = box_index(L, i, j, k);
    x[2 * v] = r;
    x[2 * v + 1] = d;
}
__global__ void k_warm_apply(LevelDesc L, const double *__restrict__ sum, double n, const float *__restrict__ y,
                             const double *__restrict__ p_prev, double *__restrict__ p) {
    int i, j, k;
    if (!inner_at(L, i, j, k)) return;
    if (!(*sum > 0.0)) return;
    const long long idx = dev::lin(L, i, j, k);
    double v = p[idx] - p_prev[idx];
    if (y != nullptr) v += sqrt(*sum / n) * static_cast<double>(y[box_index(L, i, j, k)]);
    p[idx] += v;
}
__global__ void k_warm_target(LevelDesc L, const double *__restrict__ sum, double n,
                              const double *__restrict__ p_new, const double *__restrict__ p_start,
                              const double *__restrict__ p_prev, float *__restrict__ t) {
    const int i = 1 + static_cast<int>(blockIdx.x) * BX + static_cast<int>(threadIdx.x);
    const int j = 1 + static_cast<int>(blockIdx.y) * BY + static_cast<int>(threadIdx.y);
    const int k = 1 + static_cast<int>(blockIdx.z);
    if (i > L.n[0] || j > L.n[1]) return;
    float a = 0.f;
    if (!in_obstacle(L, i, j, k)) {
        const long long idx = dev::lin(L, i, j, k);
        const double e = (p_new[idx] - p_start[idx]) - (p_start[idx] - p_prev[idx]);
        a = static_cast<float>(e * inv_rms(sum, n));
    }
    t[box_index(L, i, j, k)] = a;
}

// ---- projection start: the new increment D = p - p_prev and its image
// W = lap(p) - lap(p_prev) (both arrays carry their fill, which is affine, so
// W is the image of D under the linear part of the operator) into slot
// `fresh` of the ring; partials of (D, W_s) and (D_s, r) for every valid
// slot s, r = b - lap(p) the residual of p for the new right-hand side
__global__ void k_proj_update(LevelDesc L, const double *__restrict__ p, const double *__restrict__ p_prev,
                              const double *__restrict__ b, ProjRing ring, int fresh, unsigned valid,
                              double *__restrict__ partial) {
    int i, j, k;
    double v[2 * kProjMax];
#pragma unroll
    for (int q = 0; q < 2 * kProjMax; q++) v[q] = 0.0;
    if (inner_at(L, i, j, k)) {
        const long long idx = dev::lin(L, i, j, k);
        const double lp = dev::laplace(L, p, idx);
        const double w = lp - dev::laplace(L, p_prev, idx);
        const double r = b[idx] - lp;
        const double d = p[idx] - p_prev[idx];
        ring.d[fresh][idx] = d;
        ring.w[fresh][idx] = w;
#pragma unroll
        for (int q = 0; q < kProjMax; q++) {
            if (!((valid >> q) & 1u)) continue;
            v[q] = d * (q == fresh ? w : ring.w[q][idx]);
            v[kProjMax + q] = (q == fresh ? d : ring.d[q][idx]) * r;
        }
    }
    write_partials<2 * kProjMax>(v, partial);
}
// the sums, row and column `fresh` of H = -(D_s, W_t) (slots; the operator
// is symmetric and negative definite, so H is symmetric and positive
// definite on independent increments), and the coefficients alpha (slots)
// of the Galerkin condition (D_s, r - sum_t alpha_t W_t) = 0 over the valid
// slots, newest first, by Cholesky: the combination closest to the solution
// in the energy norm.  A slot whose increment is within 1e-5 (relative, in
// that norm) of the span of the newer ones gets no coefficient.
__global__ void k_proj_solve(const double *__restrict__ partial, int blocks, int fresh, unsigned valid,
                             double *__restrict__ G, double *__restrict__ alpha) {
    __shared__ double part[2 * kProjMax][16];
    double acc[2 * kProjMax];
#pragma unroll
    for (int q = 0; q < 2 * kProjMax; q++) acc[q] = 0.0;
    for (int t = threadIdx.x; t < blocks; t += 1024)
#pragma unroll
        for (int q = 0; q < 2 * kProjMax; q++) acc[q] += partial[q * blocks + t];
#pragma unroll
    for (int q = 0; q < 2 * kProjMax; q++) {
        double x = acc[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) part[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    // (the one thread's work arrays live in LDS: indexed at run time, private
    // arrays would go to per-lane scratch memory)
    __shared__ double g[kProjMax], y[kProjMax], Lc[kProjMax][kProjMax];
    __shared__ int order[kProjMax];
    __shared__ bool keep[kProjMax];
    for (int q = 0; q < 2 * kProjMax; q++) {
        double x = 0.0;
        for (int w = 0; w < 16; w++) x += part[q][w];
        if (q < kProjMax) {
            G[fresh * kProjMax + q] = -x;
            G[q * kProjMax + fresh] = -x;
        } else {
            g[q - kProjMax] = -x;
        }
    }
    // valid slots by age, newest first
    int m = 0;
    for (int a = 0; a < kProjMax; a++) {
        const int q = (fresh - a + kProjMax) % kProjMax;
        if ((valid >> q) & 1u) order[m++] = q;
    }
    for (int a = 0; a < m; a++) {
        double dsum = G[order[a] * kProjMax + order[a]];
        for (int c = 0; c < a; c++) dsum -= Lc[a][c] * Lc[a][c];
        keep[a] = dsum > 1e-10 * G[order[a] * kProjMax + order[a]] && dsum > 0.0;
        const double piv = keep[a] ? sqrt(dsum) : 0.0;
        Lc[a][a] = piv;
        for (int e = a + 1; e < m; e++) {
            double x = G[order[e] * kProjMax + order[a]];
            for (int c = 0; c < a; c++) x -= Lc[e][c] * Lc[a][c];
            Lc[e][a] = keep[a] ? x / piv : 0.0;
        }
    }
    for (int a = 0; a < m; a++) {
        double x = g[order[a]];
        for (int c = 0; c < a; c++) x -= Lc[a][c] * y[c];
        y[a] = keep[a] ? x / Lc[a][a] : 0.0;
    }
    for (int q = 0; q < kProjMax; q++) alpha[q] = 0.0;
    for (int a = m - 1; a >= 0; a--) {
        double x = y[a];
        for (int c = a + 1; c < m; c++) x -= Lc[c][a] * alpha[order[c]];
        alpha[order[a]] = keep[a] ? x / Lc[a][a] : 0.0;
    }
}
// p += sum over the valid slots (newest first) of alpha_s D_s
__global__ void k_proj_apply(LevelDesc L, ProjRing ring, int fresh, unsigned valid, const double *__restrict__ alpha,
                             double *__restrict__ p) {
    int i, j, k;
    if (!inner_at(L, i, j, k)) return;
    const long long idx = dev::lin(L, i, j, k);
    double x = p[idx];
#pragma unroll
    for (int a = 0; a < kProjMax; a++) {
        const int q = (fresh - a + kProjMax) % kProjMax;
        if ((valid >> q) & 1u) x += alpha[q] * ring.d[q][idx];
    }
    p[idx] = x;
}

}  // namespace

void warm_sum(hipStream_t s, const LevelDesc &L, const double *p, const double *p_prev, double *partial,
              double *sum) {
    k_warm_partials<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, p, p_prev, partial);
    k_sum<<<1, 1024, 0, s>>>(partial, blocks_inner(L), sum);
}
void warm_pack(hipStream_t s, const LevelDesc &L, const double *p, const double *b, const double *p_prev,
               const double *sum, double fluid_cells, double c, float *x) {
    k_warm_pack<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, p, b, p_prev, sum, fluid_cells, c, x);
}
void proj_update(hipStream_t s, const LevelDesc &L, const double *p, const double *p_prev, const double *b,
                 const ProjRing &ring, int fresh, unsigned valid, double *partial) {
    k_proj_update<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, p, p_prev, b, ring, fresh, valid, partial);
}
void proj_solve(hipStream_t s, const double *partial, int blocks, int fresh, unsigned valid, double *G,
                double *alpha) {
    k_proj_solve<<<1, 1024, 0, s>>>(partial, blocks, fresh, valid, G, alpha);
}
void proj_apply(hipStream_t s, const LevelDesc &L, const ProjRing &ring, int fresh, unsigned valid,
                const double *alpha, double *p) {
    k_proj_apply<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, ring, fresh, valid, alpha, p);
}
void warm_apply(hipStream_t s, const LevelDesc &L, const double *sum, double fluid_cells, const float *y,
                const double *p_prev, double *p) {
    k_warm_apply<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, sum, fluid_cells, y, p_prev, p);
}
void warm_target(hipStream_t s, const LevelDesc &L, const double *sum, double fluid_cells, const double *p_new,
                 const double *p_start, const double *p_prev, float *t) {
    k_warm_target<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, sum, fluid_cells, p_new, p_start, p_prev, t);
}

int blocks_inner(const LevelDesc &L) {
    const dim3 g = inner_grid(L);
    return static_cast<int>(g.x * g.y * g.z);
}
int blocks_dots(const LevelDesc &L) {
    const dim3 g = dots_grid(L);
    return static_cast<int>(g.x * g.y * g.z);
}
int blocks_full(const LevelDesc &L) {
    const dim3 g = update_grid(L);
    return static_cast<int>(g.x * g.y * g.z);
}

void residual_vec(hipStream_t s, const LevelDesc &L, double *r, const double *x, const double *b, double *partial) {
    k_residual_vec<<<inner_grid(L), dim3(BX, BY), 0, s>>>(L, r, x, b, partial);
}
template <typename Tz>
void lap_dots(hipStream_t s, const LevelDesc &L, const Tz *z, const Tz *zp, const double *r, double *partial) {
    k_lap_dots<Tz><<<dots_grid(L), dim3(BX, BY), 0, s>>>(L, z, zp, r, partial);
}
template void lap_dots<double>(hipStream_t, const LevelDesc &, const double *, const double *, const double *,
                               double *);
template void lap_dots<float>(hipStream_t, const LevelDesc &, const float *, const float *, const double *, double *);
void gcr_coef(hipStream_t s, const double *partial, int blocks, double *coef, const GcrGate *gate) {
    k_gcr_coef<<<1, 1024, 0, s>>>(partial, blocks, coef, gate);
}
template <typename Tz>
void gcr_update(hipStream_t s, const LevelDesc &L, const double *coef, double *x, double *r, const Tz *z,
                const Tz *zp, Tz *zp_out, double *partial, const GcrGate *gate) {
    k_gcr_update<Tz><<<update_grid(L), dim3(BX, BY), 0, s>>>(L, coef, x, r, z, zp, zp_out, partial, gate);
}
template void gcr_update<double>(hipStream_t, const LevelDesc &, const double *, double *, double *, const double *,
                                 const double *, double *, double *, const GcrGate *);
template void gcr_update<float>(hipStream_t, const LevelDesc &, const double *, double *, double *, const float *,
                                const float *, float *, double *, const GcrGate *);
void sum(hipStream_t s, const double *partial, int n, double *out) { k_sum<<<1, 1024, 0, s>>>(partial, n, out); }
void sum_gate(hipStream_t s, const double *partial, int n, GcrGate *gate) {
    k_sum_gate<<<1, 1024, 0, s>>>(partial, n, gate);
}
void gate_open(hipStream_t s, GcrGate *gate, double tol2) { k_gate_open<<<1, 64, 0, s>>>(gate, tol2); }
void gate_start(hipStream_t s, const double *partial, int n, GcrGate *gate, double tol2) {
    k_gate_start<<<1, 1024, 0, s>>>(partial, n, gate, tol2);
}
void gate_confirm(hipStream_t s, const double *partial, int n, const GcrGate *gate, double tol, LoopLog *log) {
    k_gate_confirm<<<1, 1024, 0, s>>>(partial, n, gate, tol, log);
}

}  // namespace kk
}  // namespace artss
