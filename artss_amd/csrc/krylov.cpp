#include "krylov.h"

#include <algorithm>
#include <cmath>

#include "krylov_kernels.h"

namespace artss {

GcrPressure::GcrPressure(Domain &d, VCycleMG &preconditioner, bool mixed)
    : m_domain(d), m_pre(preconditioner), m_mixed(mixed), m_r(d, P), m_z(d, P),
      m_zp0(d, P), m_zp1(d, P) {
    if (m_mixed && !m_pre.mixed_supported())
        throw config_error("the mixed-precision pressure solve needs the weighted-Jacobi smoother, boundary fills "
                           "without chained obstacle faces and a coarse-level kernel starting below level 0");
    const LevelDesc &L = d.level(0);
    const int blocks = std::max(kk::blocks_inner(L), kk::blocks_full(L));
    ARTSS_HIP(hipMalloc(&m_partial, sizeof(double) * 4 * static_cast<size_t>(blocks)));
    ARTSS_HIP(hipMalloc(&m_coef, sizeof(double) * 4));
    ARTSS_HIP(hipMalloc(&m_sum, sizeof(double)));
    ARTSS_HIP(hipHostMalloc(&m_sum_host, sizeof(double)));
    ARTSS_HIP(hipMalloc(&m_gate, sizeof(kk::GcrGate)));
    ARTSS_HIP(hipHostMalloc(&m_gate_host, sizeof(kk::GcrGate)));
    for (float *&b : m_zp32) {
        ARTSS_HIP(hipMalloc(&b, sizeof(float) * static_cast<size_t>(L.size)));
        ARTSS_HIP(hipMemset(b, 0, sizeof(float) * static_cast<size_t>(L.size)));
    }
}

GcrPressure::~GcrPressure() {
    (void)hipStreamSynchronize(m_domain.stream());
    for (auto &row : m_graph)
        for (hipGraphExec_t g : row)
            if (g) (void)hipGraphExecDestroy(g);
    (void)hipFree(m_partial);
    (void)hipFree(m_coef);
    (void)hipFree(m_sum);
    (void)hipHostFree(m_sum_host);
    (void)hipFree(m_gate);
    (void)hipHostFree(m_gate_host);
    for (float *b : m_zp32) (void)hipFree(b);
}

// One iteration: z = M r (one V-cycle from zero), the linear part of the
// fill on z, w = A z and wp = A zp with the four dot products, the
// coefficients, and the update of p, r and the kept direction (read from
// buffer `parity`, written to the other), ending with (r, r) in the gate.
// The coefficient, update and sum kernels do nothing once the gate is
// closed; the preconditioner and lap_dots write only scratch buffers.
void GcrPressure::iteration(Field &p, int parity) {
    Domain &d = m_domain;
    const LevelDesc &L = d.level(0);
    hipStream_t s = d.stream();
    if (m_mixed) {
        // the FP32 result already carries the linear part of the fill; the
        // kept direction is stored in FP32 as well
        m_pre.vcycle_mixed(m_r.data);
        const float *z = m_pre.mixed_result();
        const float *zp = m_zp32[parity];
        float *zp_out = m_zp32[parity ^ 1];
        d.timed("gcr_lap", 0, [&] { kk::lap_dots(s, L, z, zp, m_r.data, m_partial); });
        kk::gcr_coef(s, m_partial, kk::blocks_dots(L), m_coef, m_gate);
        d.timed("gcr_update", 0, [&] {
            kk::gcr_update(s, L, m_coef, p.data, m_r.data, z, zp, zp_out, m_partial, m_gate);
        });
    } else {
        m_z.set_value(0.0);
        m_pre.vcycle_eager(m_z, m_r);
        d.apply_boundary_linear(m_z.data, P, 0);
        const double *zp = (parity ? m_zp1 : m_zp0).data;
        double *zp_out = (parity ? m_zp0 : m_zp1).data;
        d.timed("gcr_lap", 0, [&] { kk::lap_dots(s, L, m_z.data, zp, m_r.data, m_partial); });
        kk::gcr_coef(s, m_partial, kk::blocks_dots(L), m_coef, m_gate);
        d.timed("gcr_update", 0, [&] {
            kk::gcr_update(s, L, m_coef, p.data, m_r.data, m_z.data, zp, zp_out, m_partial, m_gate);
        });
    }
    kk::sum_gate(s, m_partial, kk::blocks_full(L), m_gate);
}

// `count` iterations (1 or 2) as one replay of a graph captured once per
// (p, b) pair and parity of the kept direction's buffer
void GcrPressure::launch_iteration(Field &p, const Field &b, int count) {
    Domain &d = m_domain;
    const int parity = m_parity;
    m_parity = (m_parity + count) & 1;
    if (!m_pre.graphable()) {
        for (int q = 0; q < count; q++) iteration(p, (parity + q) & 1);
        return;
    }
    hipGraphExec_t &g = m_graph[count - 1][parity];
    if (!g || m_graph_p != p.data || m_graph_b != b.data) {
        if (m_graph_p != p.data || m_graph_b != b.data) {
            ARTSS_HIP(hipStreamSynchronize(d.stream()));
            for (auto &row : m_graph)
                for (hipGraphExec_t &e : row)
                    if (e) {
                        ARTSS_HIP(hipGraphExecDestroy(e));
                        e = nullptr;
                    }
            m_graph_p = p.data;
            m_graph_b = b.data;
        }
        ARTSS_HIP(hipStreamBeginCapture(d.stream(), hipStreamCaptureModeRelaxed));
        try {
            for (int q = 0; q < count; q++) iteration(p, (parity + q) & 1);
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(d.stream(), &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        hipGraph_t graph = nullptr;
        ARTSS_HIP(hipStreamEndCapture(d.stream(), &graph));
        const hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        ARTSS_HIP(e);
    }
    ARTSS_HIP(hipGraphLaunch(g, d.stream()));
}

// the reference's residual on p with its boundary fill applied
double GcrPressure::true_residual(Field &p, const Field &b) {
    m_domain.apply_boundary(p.data, P, 0);
    return m_pre.residual(p, b);
}

int GcrPressure::solve(Field &p, const Field &b, double tol, int max_iter, double *residual) {
    Domain &d = m_domain;
    if (m_mixed) m_pre.prepare_mixed();
    else m_pre.prepare_direct();
    const LevelDesc &L = d.level(0);
    hipStream_t s = d.stream();
    // convergence is rr <= tol2 on the host and in the gate alike
    const double tol2 = tol * tol;
    StepCapture &cap = d.capture();
    if (cap.mode == StepCapture::kCapture) {
        // a fixed budget of iterations behind the gate, then the true
        // residual; the outcome stays on the device
        const int slot = cap.next++;
        if (slot >= StepCapture::kMaxLoops || slot >= static_cast<int>(cap.budgets.size()))
            throw capture_unsupported("more residual-target loops than the recorded step had");
        d.apply_boundary(p.data, P, 0);
        kk::residual_vec(s, L, m_r.data, p.data, b.data, m_partial);
        if (m_mixed) k::fill32(s, m_zp32[0], 0.0f, L.size);
        else m_zp0.set_value(0.0);
        kk::gate_start(s, m_partial, kk::blocks_inner(L), m_gate, tol2);
        const int budget = std::min(cap.budgets[static_cast<size_t>(slot)], max_iter);
        for (int q = 0; q < budget; q++) iteration(p, q & 1);
        d.apply_boundary(p.data, P, 0);
        k::residual_partials(s, L, p.data, b.data, d.partial_buffer());
        kk::gate_confirm(s, d.partial_buffer(), k::partial_count(L), m_gate, tol, cap.log + slot);
        *residual = 0.0;
        return 0;
    }
    auto restart = [&]() {
        d.apply_boundary(p.data, P, 0);
        kk::residual_vec(s, L, m_r.data, p.data, b.data, m_partial);
        kk::sum(s, m_partial, kk::blocks_inner(L), m_sum);
        ARTSS_HIP(hipMemcpyAsync(m_sum_host, m_sum, sizeof(double), hipMemcpyDeviceToHost, s));
        // the kept direction starts at zero in buffer 0 (the update's first
        // write goes to buffer 1, so only buffer 0 must be cleared)
        if (m_mixed) k::fill32(s, m_zp32[0], 0.0f, L.size);
        else m_zp0.set_value(0.0);
        m_parity = 0;
        kk::gate_open(s, m_gate, tol2);
        ARTSS_HIP(hipStreamSynchronize(s));
        return *m_sum_host;
    };
    double rr = restart();
    m_initial = std::sqrt(rr);
    m_stagnated = false;
    int it = 0, it_base = 0;
    double best = rr;
    int it_best = 0;
    // right after a restart the recursive residual is the true one, up to
    // the summation order; at least one iteration follows, so every restart
    // costs an iteration and the loop ends at max_iter
    bool restarted = false;
    int restarts = 0;
    while (true) {
        if (rr <= tol2 && !restarted) {
            const double t = true_residual(p, b);
            if (t <= tol) {
                *residual = t;
                m_converged = true;
                m_predicted = std::max(0, it - 1);
                if (cap.mode == StepCapture::kRecord) {
                    // a solve that restarted cannot be replayed as one budget
                    if (restarts > 0) cap.failed = true;
                    cap.counts.push_back(it);
                    cap.kinds.push_back(StepCapture::kPressure);
                }
                return it;
            }
            // the recursive residual drifted below the true one: restart
            // from the true residual and keep iterating
            rr = restart();
            it_base = it;
            restarted = true;
            restarts++;
            continue;
        }
        if (it >= max_iter) break;
        const int batch = it < m_predicted ? std::min(m_predicted - it, max_iter - it) : 1;
        for (int q = 0; q < batch; q += 2) launch_iteration(p, b, batch - q >= 2 ? 2 : 1);
        ARTSS_HIP(hipMemcpyAsync(m_gate_host, m_gate, sizeof(kk::GcrGate), hipMemcpyDeviceToHost, s));
        ARTSS_HIP(hipStreamSynchronize(s));
        rr = m_gate_host->rr;
        it = it_base + m_gate_host->count;
        restarted = false;
        if (rr < 0.0625 * best) {
            best = rr;
            it_best = it;
        } else if (rr > tol2 && it - it_best >= kStallIterations) {
            m_stagnated = true;
            break;
        }
    }
    // a solve that ran out of iterations or stalled predicts nothing about
    // the next one: the previous prediction stays
    *residual = true_residual(p, b);
    m_converged = *residual <= tol;
    if (m_converged) m_predicted = std::max(0, it - 1);
    if (cap.mode == StepCapture::kRecord) cap.failed = true;  // ended by the iteration limit or a stall
    return it;
}

}  // namespace artss
