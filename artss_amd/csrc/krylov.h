// Krylov-accelerated pressure solve: GCR with one previous direction,
// right-preconditioned by one V-cycle of VCycleMG per iteration.
//
// Not a reference algorithm: the reference iterates V-cycles on their own
// (VCycleMG::pressure, VCycleMG.cpp:142-207).  With obstacles its V-cycle
// settles at a contraction near 0.9 per cycle after the first few cycles;
// the Krylov step removes the few slowly converging error components, so
// the same residual target ||rhs - lap p||_2 (VCycleMG.cpp:186-195) is met
// in a handful of V-cycles.  Stopping uses the recursive residual and is
// confirmed on the true residual with the reference's definition.
#pragma once

#include <memory>

#include "field.h"
#include "krylov_kernels.h"
#include "vcycle.h"

namespace artss {

class GcrPressure {
  public:
    // mixed: the preconditioner's V-cycle runs its upper levels in FP32
    // (VCycleMG::vcycle_mixed); the iteration, residuals and norms stay FP64.
    // The preconditioner's coarse solve is set on the V-cycle
    // (VCycleMG::set_precond_coarse); it need not solve its coarse problem
    // to the reference's tolerance because GCR corrects what it misses.
    GcrPressure(Domain &d, VCycleMG &preconditioner, bool mixed = false);
    ~GcrPressure();
    GcrPressure(const GcrPressure &) = delete;
    GcrPressure &operator=(const GcrPressure &) = delete;

    // iterate until the true residual ||b - lap p||_2 <= tol, or until
    // max_iter iterations; returns the iterations (V-cycles) taken and in
    // *residual the final true residual.  converged() tells the two apart:
    // a solve is converged only when its true residual meets tol.
    // Up to one fewer iterations than the previous solve took are enqueued
    // at once behind the convergence gate (kk::GcrGate), the rest one at a
    // time: the result is the same as with a check after every iteration.
    int solve(Field &p, const Field &b, double tol, int max_iter, double *residual);
    // ||b - lap p||_2 at the start of the last solve
    double initial_residual() const { return m_initial; }
    // the last solve ended with its true residual <= tol
    bool converged() const { return m_converged; }
    // the last solve stopped early because its residual stopped falling:
    // kStallIterations iterations without a fall to a quarter of the best
    // residual seen (a residual the preconditioner cannot reduce)
    bool stagnated() const { return m_stagnated; }
    static constexpr int kStallIterations = 40;

  private:
    void iteration(Field &p, int parity);  // enqueue one GCR iteration
    void launch_iteration(Field &p, const Field &b, int count = 1);
    double true_residual(Field &p, const Field &b);

    Domain &m_domain;
    VCycleMG &m_pre;
    bool m_mixed;
    Field m_r, m_z;
    // the kept direction, in two buffers that alternate between iterations
    // (the update reads one with a stencil and writes the other); FP64 for
    // the FP64 iteration, FP32 for the mixed one
    Field m_zp0, m_zp1;
    float *m_zp32[2] = {nullptr, nullptr};
    int m_parity = 0;                  // buffer holding the kept direction
    double *m_partial = nullptr;       // 4 sums x blocks
    double *m_coef = nullptr;          // device: beta, 1/|w'|, alpha
    double *m_sum = nullptr;           // device scalar
    double *m_sum_host = nullptr;      // pinned
    kk::GcrGate *m_gate = nullptr;     // device
    kk::GcrGate *m_gate_host = nullptr;  // pinned
    int m_predicted = 0;               // iterations enqueued at once in the next solve
    double m_initial = -1.0;
    bool m_converged = false;
    bool m_stagnated = false;
    // captured iterations: [count - 1][parity of the kept direction's buffer]
    hipGraphExec_t m_graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    double *m_graph_p = nullptr;
    const double *m_graph_b = nullptr;
};

}  // namespace artss
