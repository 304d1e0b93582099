#include "vcycle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mg_tail.h"

namespace artss {

VCycleMG::VCycleMG(Domain &d, const MultigridSettings &s) : IPressure(d), m_s(s) {
    if (s.smoother.type == "Jacobi") m_cgs = false;
    else if (s.smoother.type == "ColoredGaussSeidel") m_cgs = true;
    else throw config_error("multigrid smoother '" + s.smoother.type + "' is not implemented");
    if (s.n_level > d.levels()) throw config_error("pressure n_level exceeds the domain's multigrid levels");
    const int L = s.n_level;
    m_res0.resize(L + 1);
    m_res1.resize(L + 1);
    m_err0.resize(L + 1);
    m_err1.resize(L + 1);
    m_tmp.resize(L + 1);
    m_err1[0] = std::make_unique<Field>(d, P, 0.0, 0);
    m_tmp[0] = std::make_unique<Field>(d, P, 0.0, 0);
    for (int l = 0; l < L; l++) {
        m_res0[l] = std::make_unique<Field>(d, P, 0.0, l);
        m_res1[l + 1] = std::make_unique<Field>(d, P, 0.0, l + 1);
        m_err0[l] = std::make_unique<Field>(d, P, 0.0, l);
        m_err1[l + 1] = std::make_unique<Field>(d, P, 0.0, l + 1);
        m_tmp[l + 1] = std::make_unique<Field>(d, P, 0.0, l + 1);
    }
    if (d.device() >= 0) build_tail();
}

VCycleMG::~VCycleMG() {
    drop_graphs();
    free_mixed();
    free_direct();
    if (m_plan_dev) (void)hipFree(m_plan_dev);
}

bool VCycleMG::mixed_supported() const {
    if (!m_fused || m_tail_first < 1 || m_tail_first >= m_s.n_level) return false;
    for (int l = 0; l < m_tail_first; l++)
        if (!fast_level(l)) return false;
    return true;
}

void VCycleMG::free_mixed() {
    for (auto *v : {&m_err1f, &m_tmpf, &m_scrf, &m_res0f, &m_res1f}) {
        for (float *p : *v)
            if (p) (void)hipFree(p);
        v->clear();
    }
}

void VCycleMG::alloc_mixed() {
    if (!m_err1f.empty()) return;
    const int f = m_tail_first;
    // pick the cells per thread of the FP32 levels' kernels by measurement
    for (int l = 0; l < f; l++) k::tune_level(m_domain.stream(), m_domain.level(l), m_domain.bc_view(l, P, true));
    auto alloc = [&](int level) {
        float *p = nullptr;
        const size_t bytes = sizeof(float) * static_cast<size_t>(m_domain.level(level).size);
        ARTSS_HIP(hipMalloc(&p, bytes));
        ARTSS_HIP(hipMemset(p, 0, bytes));
        return p;
    };
    for (int l = 0; l < f; l++) {
        m_err1f.push_back(alloc(l));
        m_tmpf.push_back(alloc(l));
        m_scrf.push_back(alloc(l));
        m_res0f.push_back(alloc(l));
        m_res1f.push_back(l > 0 ? alloc(l) : nullptr);
    }
}

// smooth_fast in FP32: tmpf -> err1f -> ... -> err1f, linear part of the fill
// `sweeps` weighted-Jacobi sweeps from tmp into err1 (0: the XML's
// effective count).  An even count starts with one sweep into a third
// buffer, so the result always lands in err1; the first sweep into a buffer
// copies the cells it does not update, so err1 holds tmp's values there.
int VCycleMG::mixed_sweep_count(int sweeps) const {
    const int nx = m_s.n_relax;
    return sweeps > 0 ? sweeps : ((nx % 2 == 0 && nx > 0) ? nx - 1 : nx);
}

float *VCycleMG::first_sweep_target(int level, int sweeps) const {
    const int n = mixed_sweep_count(sweeps);
    return (n % 2 == 0 && n > 0) ? m_scrf[level] : m_err1f[level];
}

void VCycleMG::smooth_fast_f(int level, const double *b64, const float *b32, int sweeps, bool from_zero,
                             bool first_done) {
    Domain &d = m_domain;
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    const LevelDesc &Ld = d.level(level);
    const BCView V = d.bc_view(level, P, true);
    const int n = mixed_sweep_count(sweeps);
    const double w = m_mixed_weight > 0.0 ? m_mixed_weight : m_s.smoother.w;
    auto sweep = [&](bool first, float *dst, const float *src) {
        d.timed("jacobi_mg32", level, [&] {
            if (b64) k::jacobi_bc<float, double>(d.stream(), Ld, V, first, dst, src, b64, ax, ay, az, beta, kDsign, w);
            else k::jacobi_bc<float, float>(d.stream(), Ld, V, first, dst, src, b32, ax, ay, az, beta, kDsign, w);
        });
    };
    // from_zero: tmp is zero and the first sweep does not read it; after
    // the first sweep tmp is free and the sweeps alternate err1 and tmp.
    // first_done: the first sweep's result is already in
    // first_sweep_target(level) (restrict_sweep0)
    float *src = from_zero ? nullptr : m_tmpf[level];
    int m = n;
    if (n % 2 == 0 && n > 0) {
        if (!first_done) sweep(true, m_scrf[level], src);
        src = m_scrf[level];
        m = n - 1;
        first_done = false;
    }
    float *dst = m_err1f[level];
    for (int i = 0; i < m; i++) {
        if (!(i == 0 && first_done)) sweep(i == 0, dst, src);
        src = dst;
        dst = dst == m_err1f[level] ? m_tmpf[level] : m_err1f[level];
    }
}

namespace {

// In-place LU with partial pivoting of the n x n row-major A, then its
// inverse by n forward/back substitutions.  false when a pivot falls below
// 1e-12 of the largest entry (singular for this purpose).
bool invert_dense(std::vector<double> &A, int n, std::vector<double> &inv) {
    const size_t N = static_cast<size_t>(n);
    double scale = 0.0;
    for (double v : A) scale = std::max(scale, std::fabs(v));
    if (!(scale > 0.0)) return false;
    std::vector<int> piv(N);
    for (size_t k = 0; k < N; k++) {
        size_t p = k;
        for (size_t i = k + 1; i < N; i++)
            if (std::fabs(A[i * N + k]) > std::fabs(A[p * N + k])) p = i;
        if (!(std::fabs(A[p * N + k]) > 1e-12 * scale)) return false;
        piv[k] = static_cast<int>(p);
        if (p != k)
            for (size_t j = 0; j < N; j++) std::swap(A[k * N + j], A[p * N + j]);
        const double d = A[k * N + k];
        const double *rk = &A[k * N];
        for (size_t i = k + 1; i < N; i++) {
            double *ri = &A[i * N];
            const double l = ri[k] / d;
            ri[k] = l;
            for (size_t j = k + 1; j < N; j++) ri[j] -= l * rk[j];
        }
    }
    // the inverse by row operations on all its columns at once: X = P I,
    // then L^-1 and U^-1 (contiguous row updates, no dependent sums)
    inv.assign(N * N, 0.0);
    std::vector<size_t> perm(N);
    for (size_t i = 0; i < N; i++) perm[i] = i;
    for (size_t k = 0; k < N; k++) std::swap(perm[k], perm[static_cast<size_t>(piv[k])]);
    for (size_t i = 0; i < N; i++) inv[i * N + perm[i]] = 1.0;
    for (size_t k = 0; k < N; k++) {
        const double *xk = &inv[k * N];
        for (size_t i = k + 1; i < N; i++) {
            const double l = A[i * N + k];
            if (l == 0.0) continue;
            double *xi = &inv[i * N];
            for (size_t j = 0; j < N; j++) xi[j] -= l * xk[j];
        }
    }
    for (size_t k = N; k-- > 0;) {
        double *xk = &inv[k * N];
        const double d = A[k * N + k];
        for (size_t j = 0; j < N; j++) xk[j] /= d;
        for (size_t i = 0; i < k; i++) {
            const double u = A[i * N + k];
            if (u == 0.0) continue;
            double *xi = &inv[i * N];
            for (size_t j = 0; j < N; j++) xi[j] -= u * xk[j];
        }
    }
    return true;
}

}  // namespace

void VCycleMG::set_direct_coarse(bool on) {
    drop_graphs();
    free_direct();
    m_direct_enabled = on;
    m_direct_tried = false;
}

void VCycleMG::free_direct() {
    if (m_direct) (void)hipFree(m_direct);
    if (m_direct_cells) (void)hipFree(m_direct_cells);
    m_direct = nullptr;
    m_direct_cells = nullptr;
    m_direct_n = 0;
}

void VCycleMG::prepare_mixed() {
    alloc_mixed();
    prepare_direct();
}

void VCycleMG::prepare_direct() {
    if (m_direct_enabled && !m_direct_tried) {
        m_direct_tried = true;
        if (!build_direct() && !build_galerkin()) free_direct();
    }
}

// Column j of K is (e_j - J e_j) / beta on the fluid cells, J e_j one
// Jacobi sweep (weight 1, b = 0) of e_j carrying its linear fill: the sweep
// gives beta N e_j, N the neighbour sum, and the Jacobi iteration's fixed
// point solves (I / beta - N) x = dsign b.  With a constant null space
// (every row sum zero: Neumann or periodic everywhere) the mean diagonal
// times the averaging matrix is added, which leaves K^-1 b the solution
// with the mean of the cells removed as the null-space component.
bool VCycleMG::build_direct() {
    Domain &d = m_domain;
    const int f = m_tail_first;
    if (f < 1 || f > m_s.n_level || !fast_level(f)) return false;
    const std::vector<uint64_t> cells = d.inner_list(f);
    const int n = static_cast<int>(cells.size());
    if (n < 1 || n > kMaxDirectCells || n > k::kMaxDenseRows) return false;
    const LevelDesc &L = d.level(f);
    // obstacles at this level: their coarse faces make the level's operator
    // a poor stand-in for the finer ones near them, and the exact solve of it
    // (unlike the smoothed single-workgroup V-cycle) stalls the iteration
    if (static_cast<long long>(n) != static_cast<long long>(L.n[0]) * L.n[1] * L.n[2]) return false;
    const BCView V = d.bc_view(f, P, true);
    double ax, ay, az, beta;
    coeffs(f, &ax, &ay, &az, &beta);
    const size_t size = static_cast<size_t>(L.size);
    hipStream_t s = d.stream();
    double *src = nullptr, *dst = nullptr, *zero = nullptr;
    ARTSS_HIP(hipMalloc(&src, sizeof(double) * size));
    ARTSS_HIP(hipMalloc(&dst, sizeof(double) * size));
    ARTSS_HIP(hipMalloc(&zero, sizeof(double) * size));
    const size_t N = static_cast<size_t>(n);
    std::vector<double> K(N * N), col(size);
    try {
        k::fill(s, zero, 0.0, L.size);
        for (size_t j = 0; j < N; j++) {
            k::fill(s, src, 0.0, L.size);
            k::fill(s, src + cells[j], 1.0, 1);
            k::fill_bc<double>(s, L, V, src);
            k::jacobi_bc<double, double>(s, L, V, true, dst, src, zero, ax, ay, az, beta, kDsign, 1.0);
            ARTSS_HIP(hipMemcpyAsync(col.data(), dst, sizeof(double) * size, hipMemcpyDeviceToHost, s));
            ARTSS_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < N; i++) K[i * N + j] = ((i == j ? 1.0 : 0.0) - col[cells[i]]) / beta;
        }
    } catch (...) {
        (void)hipFree(src);
        (void)hipFree(dst);
        (void)hipFree(zero);
        throw;
    }
    (void)hipFree(src);
    (void)hipFree(dst);
    (void)hipFree(zero);
    double diag = 0.0, rows = 0.0;
    for (size_t i = 0; i < N; i++) {
        double r = 0.0;
        for (size_t j = 0; j < N; j++) r += K[i * N + j];
        rows = std::max(rows, std::fabs(r));
        diag += K[i * N + i];
    }
    diag /= static_cast<double>(n);
    if (rows <= 1e-9 * diag) {
        const double c = diag / static_cast<double>(n);
        for (double &v : K) v += c;
    }
    std::vector<double> inv;
    if (!invert_dense(K, n, inv)) return false;
    for (double &v : inv) v *= kDsign;
    std::vector<uint32_t> idx(N);
    for (size_t i = 0; i < N; i++) idx[i] = static_cast<uint32_t>(cells[i]);
    ARTSS_HIP(hipMalloc(&m_direct, sizeof(double) * N * N));
    ARTSS_HIP(hipMalloc(&m_direct_cells, sizeof(uint32_t) * N));
    ARTSS_HIP(hipMemcpy(m_direct, inv.data(), sizeof(double) * N * N, hipMemcpyHostToDevice));
    ARTSS_HIP(hipMemcpy(m_direct_cells, idx.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
    m_direct_n = n;
    return true;
}

// The Galerkin operator of the first coarse level f of the preconditioner,
// A_c = R lap P with P the prolongation from f to f - 1 (with the linear fill
// of both levels), lap the level f - 1 Laplacian of the residuum and R the
// eight-cell restriction, assembled column by column by running those
// kernels on unit vectors of the fluid cells of f, and inverted densely.
// Where a coarse level has obstacle cells its rediscretised operator is a
// poor stand-in for the finer ones (the obstacle grows when it is coarsened)
// and solving it exactly stalls GCR; A_c is what level f - 1's equation
// implies for a correction of the form P e, whatever the obstacle.
bool VCycleMG::build_galerkin() {
    Domain &d = m_domain;
    const int f = m_tail_first;
    if (f < 1 || f > m_s.n_level || !fast_level(f) || !fast_level(f - 1) || restrict_reads_ghost(f - 1)) return false;
    const std::vector<uint64_t> cells = d.inner_list(f);
    const int n = static_cast<int>(cells.size());
    if (n < 1 || n > kMaxDirectCells || n > k::kMaxDenseRows) return false;
    const LevelDesc &Lc = d.level(f), &Lf = d.level(f - 1);
    const BCView Vc = d.bc_view(f, P, true), Vf = d.bc_view(f - 1, P, true);
    const size_t nc = static_cast<size_t>(Lc.size), nf = static_cast<size_t>(Lf.size);
    hipStream_t s = d.stream();
    double *x = nullptr, *y = nullptr, *r = nullptr, *zero = nullptr, *c = nullptr, *junk = nullptr;
    auto release = [&] {
        for (double *q : {x, y, r, zero, c, junk})
            if (q) (void)hipFree(q);
    };
    const size_t N = static_cast<size_t>(n);
    std::vector<double> K(N * N), col(nc);
    try {
        ARTSS_HIP(hipMalloc(&x, sizeof(double) * nc));
        ARTSS_HIP(hipMalloc(&c, sizeof(double) * nc));
        ARTSS_HIP(hipMalloc(&junk, sizeof(double) * nc));
        ARTSS_HIP(hipMalloc(&y, sizeof(double) * nf));
        ARTSS_HIP(hipMalloc(&r, sizeof(double) * nf));
        ARTSS_HIP(hipMalloc(&zero, sizeof(double) * nf));
        k::fill(s, zero, 0.0, Lf.size);
        k::fill(s, r, 0.0, Lf.size);
        for (size_t j = 0; j < N; j++) {
            k::fill(s, x, 0.0, Lc.size);
            k::fill(s, x + cells[j], 1.0, 1);
            k::fill_bc<double>(s, Lc, Vc, x);
            k::prolongate_add_bc<double, double>(s, Lc, Lf, Vf, y, zero, x);
            k::residuum_faces<double, double>(s, Lf, Vf, r, y, zero);
            k::fill(s, c, 0.0, Lc.size);
            k::restrict_fill<double, double, double>(s, Lc, Lf, c, r, junk);
            ARTSS_HIP(hipMemcpyAsync(col.data(), c, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
            ARTSS_HIP(hipStreamSynchronize(s));
            // c = R (0 - lap P e_j) = -A_c e_j: K = -A_c, positive on the diagonal
            for (size_t i = 0; i < N; i++) K[i * N + j] = col[cells[i]];
        }
    } catch (...) {
        release();
        throw;
    }
    release();
    // a constant null space (Neumann or periodic everywhere): as in
    // build_direct, the mean diagonal times the averaging matrix is added
    double diag = 0.0, rows = 0.0;
    for (size_t i = 0; i < N; i++) {
        double sum = 0.0;
        for (size_t j = 0; j < N; j++) sum += K[i * N + j];
        rows = std::max(rows, std::fabs(sum));
        diag += K[i * N + i];
    }
    diag /= static_cast<double>(n);
    if (!(diag > 0.0)) return false;
    if (rows <= 1e-9 * diag) {
        const double add = diag / static_cast<double>(n);
        for (double &v : K) v += add;
    }
    std::vector<double> inv;
    if (!invert_dense(K, n, inv)) return false;
    for (double &v : inv) v = -v;  // A_c^-1 = -K^-1
    std::vector<uint32_t> idx(N);
    for (size_t i = 0; i < N; i++) idx[i] = static_cast<uint32_t>(cells[i]);
    ARTSS_HIP(hipMalloc(&m_direct, sizeof(double) * N * N));
    ARTSS_HIP(hipMalloc(&m_direct_cells, sizeof(uint32_t) * N));
    ARTSS_HIP(hipMemcpy(m_direct, inv.data(), sizeof(double) * N * N, hipMemcpyHostToDevice));
    ARTSS_HIP(hipMemcpy(m_direct_cells, idx.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
    m_direct_n = n;
    return true;
}

void VCycleMG::set_mixed_tail(int n_relax) {
    if (n_relax < 0) throw config_error("sweep counts must not be negative");
    drop_graphs();
    m_mixed_tail = n_relax;
}

void VCycleMG::set_precond_coarse(double coarse_tol, int max_iter) {
    if (max_iter < 0) throw config_error("the coarse sweep cap must not be negative");
    drop_graphs();
    m_pc_tol = coarse_tol;
    m_pc_iter = max_iter;
}

void VCycleMG::set_mixed_sweeps(int pre, int post, double weight) {
    if (pre < 0 || post < 0) throw config_error("sweep counts must not be negative");
    if (!(weight >= 0.0 && weight < 2.0)) throw config_error("the Jacobi weight must lie in [0, 2)");
    drop_graphs();
    m_mixed_pre = pre;
    m_mixed_post = post;
    m_mixed_weight = weight;
}

void VCycleMG::vcycle_mixed(const double *r) {
    if (!mixed_supported()) throw std::logic_error("mixed-precision V-cycle needs the fused fast path and the tail");
    alloc_mixed();
    Domain &d = m_domain;
    const int f = m_tail_first;
    hipStream_t s = d.stream();
    // every level's pre-smoothing starts from zero; below level 0 its first
    // sweep is fused into the restriction that produces the level's rhs
    const bool fused_first = mixed_sweep_count(m_mixed_pre) > 0;
    for (int level = 0; level < f; ++level) {
        const BCView V = d.bc_view(level, P, true);
        const double *b64 = level == 0 ? r : nullptr;
        const float *b32 = level == 0 ? nullptr : m_res1f[level];
        smooth_fast_f(level, b64, b32, m_mixed_pre, true, level > 0 && fused_first);
        d.timed("residuum32", level, [&] {
            if (b64) k::residuum_faces<float, double>(s, d.level(level), V, m_res0f[level], m_err1f[level], b64);
            else k::residuum_faces<float, float>(s, d.level(level), V, m_res0f[level], m_err1f[level], b32);
            // the restriction below reads the ghost layer: fill it
            if (restrict_reads_ghost(level)) k::fill_bc<float>(s, d.level(level), V, m_res0f[level]);
        });
        d.timed("restrict32", level, [&] {
            if (level + 1 < f && fused_first) {
                // the next level's first sweep from zero rides along
                double cx, cy, cz, cb;
                coeffs(level + 1, &cx, &cy, &cz, &cb);
                const double w = m_mixed_weight > 0.0 ? m_mixed_weight : m_s.smoother.w;
                k::restrict_sweep0(s, d.level(level + 1), d.level(level), d.bc_view(level + 1, P, true),
                                   m_res1f[level + 1], first_sweep_target(level + 1, m_mixed_pre), m_res0f[level], cx,
                                   cy, cz, cb, kDsign, w);
            } else if (level + 1 < f)
                k::restrict_fill<float, float, float>(s, d.level(level + 1), d.level(level), m_res1f[level + 1],
                                                      m_res0f[level], m_tmpf[level + 1]);
            else
                k::restrict_fill<float, double, double>(s, d.level(level + 1), d.level(level), m_res1[level + 1]->data,
                                                        m_res0f[level], m_err1[level + 1]->data);
        });
    }
    if (m_direct) {
        // restrict_fill zeroed err1[f]; the solve writes its fluid cells, the fill the rest it sets
        d.timed("coarse_direct", f, [&] {
            k::dense_apply(s, m_direct_n, m_direct, m_direct_cells, m_res1[f]->data, m_err1[f]->data);
            k::fill_bc<double>(s, d.level(f), d.bc_view(f, P, true), m_err1[f]->data);
        });
    } else {
        d.timed("mg_tail", f, [&] {
            mg_tail(s, m_plan_dev + 1, m_res1[f]->data, m_tail_lds_bytes[1], precond_coarse_iter(), m_mixed_tail,
                    m_mixed_tail > 0 ? m_mixed_weight : 0.0, true, m_pc_tol);
        });
    }
    for (int level = f; level > 0; --level) {
        const int fine = level - 1;
        const BCView V = d.bc_view(fine, P, true);
        d.timed("prolongate32", fine, [&] {
            if (level == f)
                k::prolongate_add_bc<float, double>(s, d.level(level), d.level(fine), V, m_tmpf[fine], m_err1f[fine],
                                                    m_err1[level]->data);
            else
                k::prolongate_add_bc<float, float>(s, d.level(level), d.level(fine), V, m_tmpf[fine], m_err1f[fine],
                                                   m_err1f[level]);
        });
        smooth_fast_f(fine, fine == 0 ? r : nullptr, fine == 0 ? nullptr : m_res1f[fine], m_mixed_post, false, false);
    }
}

void VCycleMG::set_schedule(bool fused, long long tail_cells) {
    drop_graphs();
    free_mixed();
    free_direct();
    m_direct_tried = false;
    m_fused = fused;
    m_tail_cells = tail_cells < 0 ? 0 : tail_cells;
    build_tail();
}

void VCycleMG::set_graphs(bool on) {
    drop_graphs();
    m_use_graphs = on;
}

void VCycleMG::drop_graphs() {
    for (CycleGraph &g : m_graph) {
        if (g.exec) {
            (void)hipStreamSynchronize(m_domain.stream());
            (void)hipGraphExecDestroy(g.exec);
        }
        g = CycleGraph();
    }
}

// Capture (once per out/b pair) and replay one V-cycle,
// optionally followed by the residual norm.  The capture records exactly
// the launches vcycle_fused issues eagerly; the buffer exchanges it makes on
// the host cancel out over one V-cycle, so every replay sees the pointers
// the capture saw.
void VCycleMG::launch_graph(int which, Field &out, const Field &b) {
    Domain &d = m_domain;
    CycleGraph &g = m_graph[which];
    if (!g.exec || g.out != out.data || g.b != b.data) {
        if (g.exec) {
            ARTSS_HIP(hipStreamSynchronize(d.stream()));
            ARTSS_HIP(hipGraphExecDestroy(g.exec));
        }
        g = CycleGraph();
        ARTSS_HIP(hipStreamBeginCapture(d.stream(), hipStreamCaptureModeRelaxed));
        try {
            vcycle_fused(out, b);
            if (which == 1) residual_async(out, b);
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(d.stream(), &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        hipGraph_t graph = nullptr;
        ARTSS_HIP(hipStreamEndCapture(d.stream(), &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        ARTSS_HIP(e);
        g.out = out.data;
        g.b = b.data;
    }
    ARTSS_HIP(hipGraphLaunch(g.exec, d.stream()));
}

// Choose the single-workgroup levels and upload their plan.  The kernel
// covers levels first..n_level with first <= n_level - 1 (it always
// includes the coarse solve); levels whose arrays exceed tail_cells, or
// whose boundary fill needs more waves than the plan holds, stay on the
// multi-workgroup path.
void VCycleMG::build_tail() {
    Domain &d = m_domain;
    const int L = m_s.n_level;
    m_tail_first = L;
    if (!m_fused || L == 0 || m_tail_cells == 0) return;
    int f = L;
    while (f > 0) {
        const int l = f - 1;
        if (d.level(l).size > m_tail_cells || d.face_waves(l, P) > kMaxTailWaves) break;
        f = l;
    }
    if (d.level(L).size > m_tail_cells || d.face_waves(L, P) > kMaxTailWaves) f = L;
    if (f > L - 1 || L - f + 1 > kMaxTailLevels) return;
    TailPlan plan;
    std::memset(&plan, 0, sizeof(plan));
    plan.first = f;
    plan.last = L;
    plan.n_relax = m_s.n_relax;
    plan.cgs = m_cgs ? 1 : 0;
    plan.w = m_s.smoother.w;
    plan.dsign = kDsign;
    plan.max_iter = m_s.smoother.max_iter;
    plan.tol = m_s.smoother.tol_res;
    const LevelDesc &Ls = d.level(L - 1);
    plan.solve_smooths = (Ls.N[0] <= 4 && Ls.N[1] <= 4) ? 1 : 0;
    for (int l = f; l <= L; l++) {
        TailLevel &T = plan.lv[l - f];
        T.L = d.level(l);
        T.bc = d.domain_bc(l, P);
        T.faces = d.faces_dev();
        T.params = d.params_dev(l, P);
        T.n_waves = d.face_waves(l, P);
        for (int w = 0; w < T.n_waves; w++) {
            T.wave_start[w] = d.face_wave_offset(l, P, w);
            T.wave_start[w + 1] = T.wave_start[w] + d.face_wave_count(l, P, w);
        }
        coeffs(l, &T.ax, &T.ay, &T.az, &T.beta);
        T.err1 = m_err1[l]->data;
        T.tmp = m_tmp[l]->data;
        T.res1 = l > 0 ? m_res1[l]->data : nullptr;
        T.res0 = l < L ? m_res0[l]->data : nullptr;
        T.err0 = l < L ? m_err0[l]->data : nullptr;
        for (int a = 0; a < 5; a++) T.lds_off[a] = -1;
    }
    // LDS placement, coarsest level first, whole levels, while they fit.
    // Plan 0 serves the reference's V-cycle; plan 1 the launches that skip
    // the coarsest level (skip_dead): there the coarsest level is never
    // touched and the level above it needs no residuum and no prolongation
    // target, so more of the finer levels fit in LDS.
    plan.io_err1 = m_err1[f]->data;
    const long long budget = mg_tail_lds_limit() / static_cast<long long>(sizeof(double));
    TailPlan plans[2] = {plan, plan};
    for (int dead = 0; dead < 2; dead++) {
        TailPlan &P = plans[dead];
        long long used = 0;
        for (int l = L; l >= f; l--) {
            TailLevel &T = P.lv[l - f];
            bool has[5] = {true, true, l < L, true, l < L};
            if (dead && l == L) continue;  // never touched
            if (dead && l == L - 1) has[2] = has[4] = false;
            long long need = 0;
            for (int a = 0; a < 5; a++)
                if (has[a]) need += d.level(l).size;
            if (used + need > budget) break;
            for (int a = 0; a < 5; a++)
                if (has[a]) {
                    T.lds_off[a] = static_cast<int>(used);
                    used += d.level(l).size;
                }
        }
        // a level in LDS needs every coarser one there too (the loop stops at the first misfit)
        P.lds_doubles = static_cast<int>(used);
        m_tail_lds_bytes[dead] = static_cast<int>(used * sizeof(double));
    }
    mg_tail_reserve_lds(std::max(m_tail_lds_bytes[0], m_tail_lds_bytes[1]));
    if (!m_plan_dev) ARTSS_HIP(hipMalloc(&m_plan_dev, 2 * sizeof(TailPlan)));
    ARTSS_HIP(hipMemcpy(m_plan_dev, plans, 2 * sizeof(TailPlan), hipMemcpyHostToDevice));
    m_tail_first = f;
}

void VCycleMG::coeffs(int level, double *ax, double *ay, double *az, double *beta) const {
    const LevelDesc &L = m_domain.level(level);
    *ax = 1. / (L.h[0] * L.h[0]);
    *ay = 1. / (L.h[1] * L.h[1]);
    *az = 1. / (L.h[2] * L.h[2]);
    const double rb = 2. * (*ax + *ay + *az);
    *beta = 1. / rb;
}

double VCycleMG::level_residual(const Field &out, const Field &b, int level) {
    const LevelDesc &L = m_domain.level(level);
    m_domain.timed("residual", level,
                   [&] { k::residual_partials(m_domain.stream(), L, out.data, b.data, m_domain.partial_buffer()); });
    return std::sqrt(m_domain.reduce(m_domain.partial_buffer(), k::partial_count(L)));
}

double VCycleMG::residual(const Field &out, const Field &b) { return level_residual(out, b, 0); }

void VCycleMG::residual_async(const Field &out, const Field &b) {
    const LevelDesc &L = m_domain.level(0);
    m_domain.timed("residual", 0,
                   [&] { k::residual_partials(m_domain.stream(), L, out.data, b.data, m_domain.partial_buffer()); });
    m_domain.reduce_async(m_domain.partial_buffer(), k::partial_count(L));
}

double VCycleMG::vcycle_residual(Field &out, const Field &b) {
    if (graphable()) {
        launch_graph(1, out, b);
        ARTSS_HIP(hipStreamSynchronize(m_domain.stream()));
        return std::sqrt(m_domain.reduced_value());
    }
    vcycle(out, b);
    return residual(out, b);
}

// VCycleMG.cpp:685-723.  The smoother swaps data pointers after each sweep;
// with an even number of sweeps the returned iterate is the one before the
// last sweep (x_{n-1}) and x_n is left in tmp.  The last sweep therefore
// never reaches the result and is skipped here; the result is the same.
void VCycleMG::smooth_jacobi(Field &out, Field &tmp, const Field &b, int level) {
    Domain &d = m_domain;
    d.apply_boundary(out.data, P, level);
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    tmp.copy_data(out);
    const int n = m_s.n_relax;
    const int effective = (n % 2 == 0 && n > 0) ? n - 1 : n;
    int swaps = 0;
    for (int i = 0; i < effective; i++) {
        d.timed("jacobi_mg", level, [&] {
            k::jacobi_step_mg(d.stream(), d.level(level), out.data, tmp.data, b.data, ax, ay, az, beta, kDsign,
                              m_s.smoother.w);
        });
        d.apply_boundary(out.data, P, level);
        Field::swap(tmp, out);
        swaps++;
    }
    // after the reference's n sweeps `out` holds x_effective in its own buffer
    if (swaps % 2 != 0) {
        // x_effective sits in tmp's buffer: swap the roles back so `out` owns it,
        // matching the reference's pointer state after n (even) or n+1 (odd) swaps
        Field::swap(tmp, out);
    }
}

// VCycleMG.cpp:725-798: Jacobi to tol_res or max_iter on the coarsest level,
// true residual after every sweep.
void VCycleMG::solve_jacobi(Field &out, Field &tmp, const Field &b, int level) {
    Domain &d = m_domain;
    d.apply_boundary(out.data, P, level);
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    tmp.copy_data(out);
    int it = 0;
    double res = 10000.;
    while (res > m_s.smoother.tol_res && it < m_s.smoother.max_iter) {
        d.timed("jacobi_mg", level, [&] {
            k::jacobi_step_mg(d.stream(), d.level(level), out.data, tmp.data, b.data, ax, ay, az, beta, kDsign,
                              m_s.smoother.w);
        });
        d.apply_boundary(out.data, P, level);
        res = level_residual(out, b, level);
        Field::swap(tmp, out);
        it++;
    }
    if (it % 2 != 0) Field::swap(tmp, out);
}

// VCycleMG.cpp:583-615
void VCycleMG::smooth_cgs(Field &out, const Field &b, int level) {
    Domain &d = m_domain;
    d.apply_boundary(out.data, P, level);
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    for (int i = 0; i < m_s.n_relax; i++) {
        for (int color = 0; color < 2; color++)
            d.timed("cgs_mg", level, [&] {
                k::cgs_color(d.stream(), d.level(level), color, out.data, b.data, ax, ay, az, beta, kDsign,
                             m_s.smoother.w, 0.0, nullptr, 0.0);
            });
        d.apply_boundary(out.data, P, level);
    }
}

// VCycleMG.cpp:617-683
void VCycleMG::solve_cgs(Field &out, const Field &b, int level) {
    Domain &d = m_domain;
    d.apply_boundary(out.data, P, level);
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    int it = 0;
    double res = 10000.;
    while (res > m_s.smoother.tol_res && it < m_s.smoother.max_iter) {
        for (int color = 0; color < 2; color++)
            d.timed("cgs_mg", level, [&] {
                k::cgs_color(d.stream(), d.level(level), color, out.data, b.data, ax, ay, az, beta, kDsign,
                             m_s.smoother.w, 0.0, nullptr, 0.0);
            });
        d.apply_boundary(out.data, P, level);
        res = level_residual(out, b, level);
        it++;
    }
}

// VCycleMG::Smooth :299-308
void VCycleMG::smooth(Field &out, Field &tmp, const Field &b, int level) {
    m_domain.apply_boundary(out.data, P, level);
    if (m_cgs) smooth_cgs(out, b, level);
    else smooth_jacobi(out, tmp, b, level);
}

// VCycleMG::Solve :562-581
void VCycleMG::solve(Field &out, Field &tmp, const Field &b, int level) {
    const LevelDesc &L = m_domain.level(out.get_level());
    if (L.N[0] <= 4 && L.N[1] <= 4) {
        smooth(out, tmp, b, level);
        return;
    }
    if (m_cgs) solve_cgs(out, b, level);
    else solve_jacobi(out, tmp, b, level);
}

void VCycleMG::vcycle(Field &out, const Field &b) {
    if (m_s.n_level == 0) {
        solve(out, *m_tmp[0], b, 0);
        return;
    }
    if (graphable()) launch_graph(0, out, b);
    else if (m_fused) vcycle_fused(out, b);
    else vcycle_plain(out, b);
}

// The V-cycle as GCR's preconditioner, launched eagerly (GcrPressure
// captures it in its own graph); the coarse kernel takes the
// preconditioner's coarse-solve settings (set_precond_coarse).
void VCycleMG::vcycle_eager(Field &out, const Field &b) {
    m_precond_call = true;
    try {
        if (m_s.n_level == 0) solve(out, *m_tmp[0], b, 0);
        else if (m_fused) vcycle_fused(out, b);
        else vcycle_plain(out, b);
    } catch (...) {
        m_precond_call = false;
        throw;
    }
    m_precond_call = false;
}

// Boundary fill of a pressure-type field on `level`: one launch when the
// fill is simple, the reference's face waves + ghosts otherwise.
void VCycleMG::bc(double *f, int level) {
    Domain &d = m_domain;
    if (d.bc_simple(level, P)) {
        d.apply_boundary_merged(f, P, level);
    } else if (d.bc_idempotent(level, P)) {
        d.timed("bc_merged", level, [&] { k::fill_bc(d.stream(), d.level(level), d.bc_view(level, P), f); });
    } else {
        d.apply_boundary(f, P, level);
    }
}

// Levels on the fully fused path: weighted-Jacobi smoother with at least one
// effective sweep and an idempotent fill.  There every sweep carries the
// fill after it, and a smoothing starts from tmp and ends in err1 (the
// effective sweep count is odd), so no buffers are exchanged.
bool VCycleMG::fast_level(int level) const {
    const int n = m_s.n_relax;
    const int effective = (n % 2 == 0 && n > 0) ? n - 1 : n;
    return !m_cgs && effective > 0 && m_domain.bc_idempotent(level, P);
}

// Restrict (VCycleMG.cpp:404-421) averages fine cells 2i-1 and 2i for the
// coarse cells i = 1..n_c.  Where 2 n_c > n_f, i.e. along an axis of one
// cell (n_c = n_f = 1: two-dimensional inputs), it reads the fine ghost
// layer, so the fill of the fine residuum must then be applied in full.
bool VCycleMG::restrict_reads_ghost(int level) const {
    const LevelDesc &fine = m_domain.level(level), &coarse = m_domain.level(level + 1);
    for (int a = 0; a < 3; a++)
        if (2 * coarse.n[a] > fine.n[a]) return true;
    return false;
}

// Smoothing of err1[level] whose initial iterate (with the fill applied)
// is in tmp[level]: sweep 1 tmp -> err1 with the copy folded in, then
// alternating, the last one into err1.
void VCycleMG::smooth_fast(int level, const double *b) {
    Domain &d = m_domain;
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    const LevelDesc &Ld = d.level(level);
    const BCView V = d.bc_view(level, P);
    const int n = m_s.n_relax;
    const int effective = (n % 2 == 0 && n > 0) ? n - 1 : n;
    double *src = m_tmp[level]->data, *dst = m_err1[level]->data;
    for (int i = 0; i < effective; i++) {
        d.timed("jacobi_mg", level, [&] {
            k::jacobi_bc(d.stream(), Ld, V, i == 0, dst, src, b, ax, ay, az, beta, kDsign, m_s.smoother.w);
        });
        std::swap(src, dst);
    }
}

// VCycleMG::Smooth on err1[level] with launches folded:
//  - the smoother's own fill right after Smooth's is dropped when the fill is
//    simple (applying it twice equals applying it once);
//  - Jacobi: the copy tmp <- out and the first sweep out <- J(tmp) become one
//    kernel writing J(out) on inner cells and out elsewhere into tmp; the
//    sweeps then alternate between the two buffers and the odd count leaves
//    the result in tmp, so the two Field objects trade buffers.  Both buffers
//    hold exactly what the reference leaves in them, with the roles exchanged.
void VCycleMG::smooth_fused(int level, const double *b) {
    Domain &d = m_domain;
    Field &out = *m_err1[level];
    Field &tmp = *m_tmp[level];
    bc(out.data, level);
    if (!d.bc_idempotent(level, P)) d.apply_boundary(out.data, P, level);
    double ax, ay, az, beta;
    coeffs(level, &ax, &ay, &az, &beta);
    const LevelDesc &Ld = d.level(level);
    if (m_cgs) {
        for (int i = 0; i < m_s.n_relax; i++) {
            for (int color = 0; color < 2; color++)
                d.timed("cgs_mg", level, [&] {
                    k::cgs_color(d.stream(), Ld, color, out.data, b, ax, ay, az, beta, kDsign, m_s.smoother.w, 0.0,
                                 nullptr, 0.0);
                });
            bc(out.data, level);
        }
        return;
    }
    const int n = m_s.n_relax;
    const int effective = (n % 2 == 0 && n > 0) ? n - 1 : n;  // odd, or 0
    if (effective == 0) return;
    d.timed("jacobi_mg", level, [&] {
        k::jacobi_mg_first(d.stream(), Ld, tmp.data, out.data, b, ax, ay, az, beta, kDsign, m_s.smoother.w);
    });
    bc(tmp.data, level);
    double *src = tmp.data, *dst = out.data;
    for (int i = 1; i < effective; i++) {
        d.timed("jacobi_mg", level, [&] {
            k::jacobi_step_mg(d.stream(), Ld, dst, src, b, ax, ay, az, beta, kDsign, m_s.smoother.w);
        });
        bc(dst, level);
        std::swap(src, dst);
    }
    Field::swap(out, tmp);  // the last sweep wrote into tmp's buffer
}

// VCycleMG::VCycleMultigrid with the fused schedule.  Besides the folds of
// smooth_fused, exact by construction:
//  - the fill of residuum1 after Restrict (:248) is dropped: only inner cells
//    of residuum1 are ever read;
//  - for a simple fill, the ghost part of the fill of residuum0 (:245) is
//    dropped where Restrict reads no ghost (every axis of the level has at
//    least two cells; see restrict_reads_ghost), Restrict and the reset of the coarse
//    error share one kernel, Prolongate + fill + correction (:266-270) become
//    one kernel adding into error1 (every cell the fill of error0 writes is
//    overwritten by the fill at the start of the next Smooth/Solve before it
//    is read), and the final fill of `out` (:283) is dropped (out already
//    satisfies it);
//  - the levels from tail_first down run in one single-workgroup kernel.
void VCycleMG::vcycle_fused(Field &out, const Field &b) {
    Domain &d = m_domain;
    const int L = m_s.n_level;
    const int f = m_tail_first;
    // the buffer the smoothing of `level` starts from
    auto start = [&](int level) -> Field & { return level < f && fast_level(level) ? *m_tmp[level] : *m_err1[level]; };
    if (f > 0 && fast_level(0)) {
        // err1 <- out and the two fills of Smooth fold into tmp <- fill(out)
        d.timed("copy_bc", 0, [&] { k::copy_bc(d.stream(), d.level(0), d.bc_view(0, P), m_tmp[0]->data, out.data); });
    } else {
        m_err1[0]->copy_data(out);
    }
    for (int level = 0; level < f; ++level) {
        const double *rhs = level == 0 ? b.data : m_res1[level]->data;
        // at levels > 0 the smoothing starts from zero, which every fill
        // (all values are zero below level 0) maps to zero: fills elided
        if (fast_level(level)) smooth_fast(level, rhs);
        else smooth_fused(level, rhs);
        if (d.bc_idempotent(level, P) && !restrict_reads_ghost(level)) {
            d.timed("residuum", level, [&] {
                k::residuum_faces(d.stream(), d.level(level), d.bc_view(level, P), m_res0[level]->data,
                                  m_err1[level]->data, rhs);
            });
        } else {
            d.timed("residuum", level, [&] {
                k::residuum(d.stream(), d.level(level), m_res0[level]->data, m_err1[level]->data, rhs);
            });
            d.apply_boundary(m_res0[level]->data, P, level);
        }
        d.timed("restrict", level, [&] {
            k::restrict_fill(d.stream(), d.level(level + 1), d.level(level), m_res1[level + 1]->data,
                             m_res0[level]->data, start(level + 1).data);
        });
    }
    if (f < L && m_precond_call && m_direct && f > 0) {
        // as GCR's preconditioner: the exact coarse solve of level f
        // (restrict_fill above zeroed err1[f])
        d.timed("coarse_direct", f, [&] {
            k::dense_apply(d.stream(), m_direct_n, m_direct, m_direct_cells, m_res1[f]->data, m_err1[f]->data);
            k::fill_bc<double>(d.stream(), d.level(f), d.bc_view(f, P, true), m_err1[f]->data);
        });
    } else if (f < L) {
        d.timed("mg_tail", f, [&] {
            if (m_precond_call)
                mg_tail(d.stream(), m_plan_dev + 1, f == 0 ? b.data : m_res1[f]->data, m_tail_lds_bytes[1],
                        precond_coarse_iter(), 0, 0.0, true, m_pc_tol);
            else
                mg_tail(d.stream(), m_plan_dev, f == 0 ? b.data : m_res1[f]->data, m_tail_lds_bytes[0]);
        });
    }
    for (int level = f; level > 0; --level) {
        const int fine = level - 1;
        const double *rhs = fine == 0 ? b.data : m_res1[fine]->data;
        if (level != L && fast_level(fine)) {
            d.timed("prolongate", fine, [&] {
                k::prolongate_add_bc(d.stream(), d.level(level), d.level(fine), d.bc_view(fine, P),
                                     m_tmp[fine]->data, m_err1[fine]->data, m_err1[level]->data);
            });
            smooth_fast(fine, rhs);
            continue;
        }
        if (d.bc_idempotent(fine, P)) {
            d.timed("prolongate", fine, [&] {
                k::prolongate_add(d.stream(), d.level(level), d.level(fine), m_err1[fine]->data,
                                  m_err1[level]->data);
            });
        } else {
            d.timed("prolongate", fine, [&] {
                k::prolongate(d.stream(), d.level(level), d.level(fine), m_err0[fine]->data, m_err1[level]->data);
            });
            d.apply_boundary(m_err0[fine]->data, P, fine);
            *m_err1[fine] += *m_err0[fine];
        }
        if (level == L) {
            Field rhs_view(d, P, fine, const_cast<double *>(rhs));
            solve(*m_err1[fine], *m_tmp[fine], rhs_view, fine);
        } else {
            smooth_fused(fine, rhs);
        }
    }
    out.copy_data(*m_err1[0]);
    if (!d.bc_idempotent(0, P)) d.apply_boundary(out.data, out.get_type(), 0);
}

// VCycleMG::VCycleMultigrid :215-288, one launch per reference operator.
// residuum1[0] is the right-hand side itself (the reference copies b there
// and never writes it), and the copy of error0[0] into out at :254 is dead
// (out is overwritten at :282).
void VCycleMG::vcycle_plain(Field &out, const Field &b) {
    Domain &d = m_domain;
    const int L = m_s.n_level;
    m_err1[0]->copy_data(out);
    for (int level = 0; level < L; ++level) {
        const Field &rhs = level == 0 ? b : *m_res1[level];
        smooth(*m_err1[level], *m_tmp[level], rhs, level);
        d.timed("residuum", level, [&] {
            k::residuum(d.stream(), d.level(level), m_res0[level]->data, m_err1[level]->data, rhs.data);
        });
        d.apply_boundary(m_res0[level]->data, P, level);
        d.timed("restrict", level, [&] {
            k::restrict_(d.stream(), d.level(level + 1), d.level(level), m_res1[level + 1]->data,
                         m_res0[level]->data);
        });
        d.apply_boundary(m_res1[level + 1]->data, P, level + 1);
        m_err1[level + 1]->set_value(0);
    }
    for (int level = L; level > 0; --level) {
        const Field &rhs = level - 1 == 0 ? b : *m_res1[level - 1];
        d.timed("prolongate", level - 1, [&] {
            k::prolongate(d.stream(), d.level(level), d.level(level - 1), m_err0[level - 1]->data,
                          m_err1[level]->data);
        });
        d.apply_boundary(m_err0[level - 1]->data, P, level - 1);
        *m_err1[level - 1] += *m_err0[level - 1];
        if (level == L) solve(*m_err1[level - 1], *m_tmp[level - 1], rhs, level - 1);
        else smooth(*m_err1[level - 1], *m_tmp[level - 1], rhs, level - 1);
    }
    out.copy_data(*m_err1[0]);
    d.apply_boundary(out.data, out.get_type(), 0);
}

void VCycleMG::pressure(Field &out, const Field &b, double t, bool sync) {
    const double dt = m_domain.settings().dt;
    const size_t Nt = static_cast<size_t>(std::round(t / dt));
    int act = 0;
    if (Nt == 1) {
        size_t set_relax = static_cast<size_t>(m_s.n_relax);
        double r = 10000.;
        while (r > m_s.tol_res && act < m_s.max_cycle && set_relax < static_cast<size_t>(m_s.smoother.max_iter)) {
            for (int i = 0; i < m_s.n_cycle; i++) {
                vcycle(out, b);
                act++;
            }
            r = residual(out, b);
            set_relax += static_cast<size_t>(m_s.n_relax);
        }
    } else {
        for (int i = 0; i < m_s.n_cycle; i++) {
            vcycle(out, b);
            act++;
        }
    }
    m_cycles = act;
    if (sync) ARTSS_HIP(hipStreamSynchronize(m_domain.stream()));
}

}  // namespace artss
