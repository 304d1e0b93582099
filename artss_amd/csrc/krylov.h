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
    // (VCycleMG::vcycle_mixed); the iteration, residuals and norms stay FP64
    // coarse_max_iter (mixed only, > 0): cap of the preconditioner's coarse
    // Jacobi solve; the preconditioner need not solve its coarse problem to
    // the reference's tolerance because GCR corrects what it misses
    GcrPressure(Domain &d, VCycleMG &preconditioner, bool mixed = false, int coarse_max_iter = 0);
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

  private:
    void iteration(Field &p);          // enqueue one GCR iteration
    void launch_iteration(Field &p, const Field &b, int count = 1);
    double true_residual(Field &p, const Field &b);

    Domain &m_domain;
    VCycleMG &m_pre;
    bool m_mixed;
    int m_coarse_max_iter;
    Field m_r, m_z, m_zp, m_wp;
    float *m_zp32 = nullptr, *m_wp32 = nullptr;  // kept direction of the mixed iteration
    double *m_partial = nullptr;       // 4 sums x blocks
    double *m_coef = nullptr;          // device: beta, 1/|w'|, alpha
    double *m_sum = nullptr;           // device scalar
    double *m_sum_host = nullptr;      // pinned
    kk::GcrGate *m_gate = nullptr;     // device
    kk::GcrGate *m_gate_host = nullptr;  // pinned
    int m_predicted = 0;               // iterations enqueued at once in the next solve
    double m_initial = -1.0;
    bool m_converged = false;
    hipGraphExec_t m_graph = nullptr, m_graph2 = nullptr;  // one and two iterations
    double *m_graph_p = nullptr;
    const double *m_graph_b = nullptr;
};

}  // namespace artss
