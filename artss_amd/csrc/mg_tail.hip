// Single-workgroup kernel for the coarse end of a multigrid V-cycle.
//
// The levels it covers hold at most a few ten thousand cells, so one
// workgroup of 1024 threads (16 wavefronts on one CU) walks each of them in
// a strided loop and a workgroup barrier separates consecutive operators.
// Every operator is the one VCycleMG applies at that point
// (VCycleMG.cpp:215-288, :299-308, :562-798), evaluated with the same
// per-cell functions as the multi-workgroup kernels (stencil.h), so the
// result is bit-identical to issuing those kernels one by one; only the
// coarse solve's residual sum is formed in this kernel's own fixed order.
#include "mg_plan.h"
#include "mg_tail.h"
#include "stencil.h"

#include <mutex>

namespace artss {
namespace {

using dev::lin;

// Each thread handles kBatch cells per pass: their loads are all issued
// before the first store, so a pass costs one memory round trip instead of
// kBatch of them.
constexpr int kBatch = 4;
constexpr int kStride = kTailThreads * kBatch;

struct Cell {
    int idx;   // linear index in the level array (levels here are small)
    bool on;   // an inner (fluid) cell
};

// inner-box cell number t (x fastest) of level L
__device__ inline Cell box_cell(const LevelDesc &L, int t, int n) {
    Cell c{0, false};
    if (t >= n) return c;
    const int nx = L.n[0], ny = L.n[1];
    const int i = 1 + t % nx, r = t / nx;
    const int j = 1 + r % ny, k = 1 + r / ny;
    c.on = !in_obstacle(L, i, j, k);
    c.idx = i + L.N[0] * (j + L.N[1] * k);
    return c;
}

__device__ inline int box_cells(const LevelDesc &L) { return L.n[0] * L.n[1] * L.n[2]; }

// DomainController::apply_boundary: obstacle faces wave by wave, then ghosts
__device__ __forceinline__ void t_bc(const TailLevel &T, double *f) {
    for (int w = 0; w < T.n_waves; w++) {
        const int e0 = static_cast<int>(T.wave_start[w]), e1 = static_cast<int>(T.wave_start[w + 1]);
        for (int e = e0 + threadIdx.x; e < e1; e += kTailThreads) dev::face_write(T.faces[e], T.params, f);
        __syncthreads();
    }
    if (T.bc.active) {
        const int n = static_cast<int>(dev::ghost_count(T.L));
        for (int t = threadIdx.x; t < n; t += kTailThreads) dev::domain_ghost(T.L, T.bc, f, t);
        __syncthreads();
    }
}

__device__ __forceinline__ void t_copy(double *__restrict__ dst, const double *__restrict__ src, int n) {
    for (int base = threadIdx.x; base < n; base += kStride) {
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = base + u * kTailThreads;
            if (t < n) v[u] = src[t];
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = base + u * kTailThreads;
            if (t < n) dst[t] = v[u];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void t_fill(double *dst, double v, int n) {
    for (int t = threadIdx.x; t < n; t += kTailThreads) dst[t] = v;
    __syncthreads();
}

__device__ __forceinline__ void t_add(double *__restrict__ dst, const double *__restrict__ src, int n) {
    for (int base = threadIdx.x; base < n; base += kStride) {
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = base + u * kTailThreads;
            if (t < n) v[u] = dst[t] + src[t];
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = base + u * kTailThreads;
            if (t < n) dst[t] = v[u];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void t_jacobi(const TailLevel &T, const TailPlan &P, double *__restrict__ out,
                         const double *__restrict__ in, const double *__restrict__ b) {
    const int n = box_cells(T.L);
    for (int base = threadIdx.x; base < n; base += kStride) {
        Cell c[kBatch];
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            c[u] = box_cell(T.L, base + u * kTailThreads, n);
            if (c[u].on) v[u] = dev::jacobi_mg(T.L, T.ax, T.ay, T.az, T.beta, P.dsign, P.w, in, b, c[u].idx);
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (c[u].on) out[c[u].idx] = v[u];
    }
    __syncthreads();
}

// one color, in place: cells of a color read only cells of the other color
__device__ __forceinline__ void t_cgs_color(const TailLevel &T, const TailPlan &P, int color, double *out, const double *b) {
    const int n = box_cells(T.L);
    const int nx = T.L.n[0], ny = T.L.n[1];
    for (int base = threadIdx.x; base < n; base += kStride) {
        Cell c[kBatch];
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = base + u * kTailThreads;
            c[u] = box_cell(T.L, t, n);
            if (c[u].on) {
                const int i = 1 + t % nx, r = t / nx;
                const int j = 1 + r % ny, k = 1 + r / ny;
                c[u].on = ((i + j + k) & 1) == color;
            }
            if (c[u].on) v[u] = dev::cgs_mg(T.L, T.ax, T.ay, T.az, T.beta, P.dsign, P.w, out, b, c[u].idx);
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (c[u].on) out[c[u].idx] = v[u];
    }
    __syncthreads();
}

__device__ __forceinline__ void t_residuum(const TailLevel &T, double *__restrict__ out, const double *__restrict__ x,
                           const double *__restrict__ b) {
    const int n = box_cells(T.L);
    for (int base = threadIdx.x; base < n; base += kStride) {
        Cell c[kBatch];
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            c[u] = box_cell(T.L, base + u * kTailThreads, n);
            if (c[u].on) v[u] = b[c[u].idx] - dev::laplace(T.L, x, c[u].idx);
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (c[u].on) out[c[u].idx] = v[u];
    }
    __syncthreads();
}

// ||b - lap x||_2 over the inner cells, same value in every thread
__device__ __forceinline__ double t_residual(const TailLevel &T, const double *__restrict__ x, const double *__restrict__ b,
                             double *red) {
    const int n = box_cells(T.L);
    double sq = 0.0;
    for (int base = threadIdx.x; base < n; base += kStride) {
        Cell c[kBatch];
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            c[u] = box_cell(T.L, base + u * kTailThreads, n);
            v[u] = c[u].on ? b[c[u].idx] - dev::laplace(T.L, x, c[u].idx) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) sq += v[u] * v[u];
    }
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kTailThreads / 64; w++) s += red[w];
        red[kTailThreads / 64] = sqrt(s);
    }
    __syncthreads();
    const double r = red[kTailThreads / 64];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void t_restrict(const TailLevel &Tc, const TailLevel &Tf, double *__restrict__ out,
                           const double *__restrict__ in) {
    const int n = box_cells(Tc.L);
    const int nx = Tc.L.n[0], ny = Tc.L.n[1];
    for (int base = threadIdx.x; base < n; base += kStride) {
        Cell c[kBatch];
        double v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int t = base + u * kTailThreads;
            c[u] = box_cell(Tc.L, t, n);
            if (c[u].on) {
                const int i = 1 + t % nx, r = t / nx;
                const int j = 1 + r % ny, k = 1 + r / ny;
                v[u] = dev::restrict_cell(Tf.L, in, i, j, k);
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (c[u].on) out[c[u].idx] = v[u];
    }
    __syncthreads();
}

__device__ __forceinline__ void t_prolongate(const TailLevel &Tc, const TailLevel &Tf, double *__restrict__ out,
                             const double *__restrict__ in) {
    const int n = box_cells(Tc.L);
    const int nx = Tc.L.n[0], ny = Tc.L.n[1];
    for (int t = threadIdx.x; t < n; t += kTailThreads) {
        const Cell c = box_cell(Tc.L, t, n);
        if (!c.on) continue;
        const int i = 1 + t % nx, r = t / nx;
        const int j = 1 + r % ny, k = 1 + r / ny;
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = dev::prolong_value(Tc.L, in, c.idx, q);
#pragma unroll
        for (int q = 0; q < 8; q++) out[dev::prolong_target(Tf.L, i, j, k, q)] = v[q];
    }
    __syncthreads();
}

// VCycleMG.cpp:685-723.  The reference swaps the two buffers after every
// sweep and once more after an odd count, so `out` always ends in its own
// buffer, which holds the last sweep for odd n_relax and the one before it
// for even n_relax.  That last sweep of an even count is never read and is
// skipped; the remaining count is odd and leaves its result in `out`.
__device__ __forceinline__ void t_smooth_jacobi(const TailLevel &T, const TailPlan &P, double *out, double *tmp, const double *b) {
    t_bc(T, out);
    t_copy(tmp, out, static_cast<int>(T.L.size));
    const int n = (P.n_relax % 2 == 0 && P.n_relax > 0) ? P.n_relax - 1 : P.n_relax;
    double *o = out, *s = tmp;
    for (int i = 0; i < n; i++) {
        t_jacobi(T, P, o, s, b);
        t_bc(T, o);
        double *x = o;
        o = s;
        s = x;
    }
}

// The coarse solve's stopping test: coarse_tol == 0, the reference's
// res > tol (absolute); coarse_tol > 0, res > coarse_tol * (the residual of
// the starting iterate), a reduction independent of the right-hand side's
// scale; coarse_tol < 0, no test (max_iter sweeps, no residual computed).
__device__ __forceinline__ double solve_target(const TailLevel &T, const TailPlan &P, double coarse_tol, const double *x,
                               const double *b, double *red) {
    if (coarse_tol == 0.0) return P.tol;
    if (coarse_tol < 0.0) return -1.0;
    return coarse_tol * t_residual(T, x, b, red);
}

// VCycleMG.cpp:725-798
__device__ __forceinline__ void t_solve_jacobi(const TailLevel &T, const TailPlan &P, double *out, double *tmp, const double *b,
                               double *red, double coarse_tol) {
    t_bc(T, out);
    t_copy(tmp, out, static_cast<int>(T.L.size));
    double *o = out, *s = tmp;
    int it = 0;
    double res = 10000.;
    const double target = solve_target(T, P, coarse_tol, o, b, red);
    while (res > target && it < P.max_iter) {
        t_jacobi(T, P, o, s, b);
        t_bc(T, o);
        if (coarse_tol >= 0.0) res = t_residual(T, o, b, red);
        double *x = o;
        o = s;
        s = x;
        it++;
    }
    // VCycleMG.cpp:793-795 swaps back after an odd count, so `out` keeps its
    // buffer, which holds x_it for odd it and x_(it-1) for even it
}

// VCycleMG.cpp:583-615
__device__ __forceinline__ void t_smooth_cgs(const TailLevel &T, const TailPlan &P, double *out, const double *b) {
    t_bc(T, out);
    for (int i = 0; i < P.n_relax; i++) {
        t_cgs_color(T, P, 0, out, b);
        t_cgs_color(T, P, 1, out, b);
        t_bc(T, out);
    }
}

// VCycleMG.cpp:617-683
__device__ __forceinline__ void t_solve_cgs(const TailLevel &T, const TailPlan &P, double *out, const double *b, double *red,
                            double coarse_tol) {
    t_bc(T, out);
    int it = 0;
    double res = 10000.;
    const double target = solve_target(T, P, coarse_tol, out, b, red);
    while (res > target && it < P.max_iter) {
        t_cgs_color(T, P, 0, out, b);
        t_cgs_color(T, P, 1, out, b);
        t_bc(T, out);
        if (coarse_tol >= 0.0) res = t_residual(T, out, b, red);
        it++;
    }
}

// VCycleMG::Smooth :299-308
__device__ __forceinline__ void t_smooth(const TailLevel &T, const TailPlan &P, double *out, double *tmp, const double *b) {
    t_bc(T, out);
    if (P.cgs) t_smooth_cgs(T, P, out, b);
    else t_smooth_jacobi(T, P, out, tmp, b);
}

// VCycleMG::Solve :562-581
__device__ __forceinline__ void t_solve(const TailLevel &T, const TailPlan &P, double *out, double *tmp, const double *b,
                        double *red, double coarse_tol) {
    if (P.solve_smooths) {
        t_smooth(T, P, out, tmp, b);
        return;
    }
    if (P.cgs) t_solve_cgs(T, P, out, b, red, coarse_tol);
    else t_solve_jacobi(T, P, out, tmp, b, red, coarse_tol);
}

__global__ void __launch_bounds__(kTailThreads) k_mg_tail(const TailPlan *__restrict__ plan, const double *b_first,
                                                           int max_iter, int n_relax, double w, int skip_dead,
                                                           double coarse_tol) {
    __shared__ TailPlan P;
    __shared__ double red[kTailThreads / 64 + 1];
    extern __shared__ double lds[];
    {
        const unsigned long long *src = reinterpret_cast<const unsigned long long *>(plan);
        unsigned long long *dst = reinterpret_cast<unsigned long long *>(&P);
        constexpr int words = sizeof(TailPlan) / sizeof(unsigned long long);
        for (int t = threadIdx.x; t < words; t += kTailThreads) dst[t] = src[t];
        __syncthreads();
        if (max_iter > 0 && threadIdx.x == 0) P.max_iter = max_iter;
        if (n_relax > 0 && threadIdx.x == 0) P.n_relax = n_relax;
        if (w > 0.0 && threadIdx.x == 0) P.w = w;
        __syncthreads();
    }
    const int f = P.first, Lm = P.last;
    // Level arrays placed in LDS start as the HBM arrays they stand for: zero
    // wherever no operator writes them (every array is zero-initialised and
    // those cells are never written), the first level's error and
    // right-hand side copied in.  Only the first level's error leaves.
    if (P.lds_doubles > 0) {
        for (int t = threadIdx.x; t < P.lds_doubles; t += kTailThreads) lds[t] = 0.0;
        if (threadIdx.x == 0) {
            for (int l = 0; l <= Lm - f; l++) {
                TailLevel &T = P.lv[l];
                double **ptr[5] = {&T.err1, &T.tmp, &T.res0, &T.res1, &T.err0};
                for (int a = 0; a < 5; a++)
                    if (T.lds_off[a] >= 0) *ptr[a] = lds + T.lds_off[a];
            }
        }
        __syncthreads();
        if (P.lv[0].lds_off[0] >= 0) {
            t_copy(P.lv[0].err1, P.io_err1, static_cast<int>(P.lv[0].L.size));
            t_copy(P.lv[0].res1, b_first, static_cast<int>(P.lv[0].L.size));
            b_first = P.lv[0].res1;
        }
    }
    // descent, VCycleMG.cpp:228-252
    for (int level = f; level < Lm; ++level) {
        const TailLevel &T = P.lv[level - f];
        const TailLevel &Tc = P.lv[level + 1 - f];
        const double *rhs = level == f ? b_first : T.res1;
        t_smooth(T, P, T.err1, T.tmp, rhs);
        // the coarsest level only ever holds a zero error (VCycleMG.cpp:228-282
        // restrict to it, then prolongate its zero back): skipped on request
        if (skip_dead && level + 1 == Lm) break;
        t_residuum(T, T.res0, T.err1, rhs);
        t_bc(T, T.res0);
        t_restrict(Tc, T, Tc.res1, T.res0);
        t_bc(Tc, Tc.res1);
        t_fill(Tc.err1, 0.0, static_cast<int>(Tc.L.size));
    }
    // ascent, VCycleMG.cpp:255-281
    for (int level = Lm; level > f; --level) {
        const TailLevel &T = P.lv[level - 1 - f];
        const TailLevel &Tc = P.lv[level - f];
        const double *rhs = level - 1 == f ? b_first : T.res1;
        if (!(skip_dead && level == Lm)) {
            t_prolongate(Tc, T, T.err0, Tc.err1);
            t_bc(T, T.err0);
            t_add(T.err1, T.err0, static_cast<int>(T.L.size));
        }
        if (level == Lm) t_solve(T, P, T.err1, T.tmp, rhs, red, coarse_tol);
        else t_smooth(T, P, T.err1, T.tmp, rhs);
    }
    if (P.lv[0].lds_off[0] >= 0) t_copy(P.io_err1, P.lv[0].err1, static_cast<int>(P.lv[0].L.size));
}

}  // namespace

int mg_tail_lds_limit() {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    const int static_bytes = static_cast<int>(sizeof(TailPlan)) + 1024;
    return bytes > static_bytes ? bytes - static_bytes : 0;
}

// The attribute is per kernel function, shared by every solver in the
// process: it is only ever raised, so a smaller plan set up later does not
// invalidate the launches of a larger one.
void mg_tail_reserve_lds(int lds_bytes) {
    static std::mutex lock;
    static int reserved = 0;
    std::lock_guard<std::mutex> guard(lock);
    if (lds_bytes > reserved &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mg_tail), hipFuncAttributeMaxDynamicSharedMemorySize,
                            lds_bytes) == hipSuccess)
        reserved = lds_bytes;
}

void mg_tail(hipStream_t s, const TailPlan *plan, const double *b_first, int lds_bytes, int max_iter, int n_relax,
             double w, bool skip_dead, double coarse_tol) {
    k_mg_tail<<<1, kTailThreads, lds_bytes, s>>>(plan, b_first, max_iter, n_relax, w, skip_dead ? 1 : 0,
                                                 coarse_tol);
}

}  // namespace artss
