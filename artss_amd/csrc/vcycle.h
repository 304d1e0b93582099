// Geometric multigrid V-cycle pressure solver (pressure/VCycleMG.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "mg_plan.h"
#include "operators.h"

namespace artss {

class VCycleMG : public IPressure {
  public:
    VCycleMG(Domain &d, const MultigridSettings &s);
    ~VCycleMG() override;

    // VCycleMG::pressure, VCycleMG.cpp:142-207: residual-driven loop in the
    // first time step (round(t/dt) == 1), otherwise n_cycle V-cycles.
    void pressure(Field &out, const Field &b, double t, bool sync) override;

    // ||b - lap_h(out)||_2 over the level-0 inner cells (VCycleMG.cpp:186-195)
    double residual(const Field &out, const Field &b);
    void vcycle(Field &out, const Field &b);
    // one V-cycle followed by the residual norm, one host round trip
    double vcycle_residual(Field &out, const Field &b);
    // one V-cycle issued launch by launch (never from a graph), for callers
    // that capture it into graphs of their own
    void vcycle_eager(Field &out, const Field &b);

    int last_cycles() const { return m_cycles; }

    // Launch schedule (results are identical either way):
    //  fused = false: one launch per reference operator, host-side coarse
    //                 solve loop (a readback per iteration);
    //  fused = true:  copies and boundary fills that cannot change the result
    //                 dropped or folded into neighbouring kernels, and every
    //                 level whose array has at most `tail_cells` cells run by
    //                 one single-workgroup kernel (0: no such kernel).
    void set_schedule(bool fused, long long tail_cells);
    bool fused() const { return m_fused; }
    int tail_first() const { return m_tail_first; }
    int tail_lds_bytes() const { return m_tail_lds_bytes[0]; }
    // With the fused schedule and the single-workgroup kernel in use, a
    // V-cycle has no host round trip; it is then captured once into a HIP
    // graph (per output/right-hand-side pair) and replayed.
    void set_graphs(bool on);
    // Mixed precision (used as the preconditioner of GcrPressure): one
    // V-cycle for A z = r from z = 0 with the linear part of the fill, the
    // levels above the single-workgroup kernel in FP32 (smoothing, residuum,
    // restriction, prolongation), the coarse levels and their solve in FP64.
    // The result stays in mixed_result() (FP32, level 0) until the next call.
    bool mixed_supported() const;
    void prepare_mixed();  // allocates (and assembles the direct solve); call before capturing vcycle_mixed
    // assembles the direct coarse solve that vcycle_eager uses as GCR's
    // preconditioner (method 1); call before capturing it
    void prepare_direct();
    // pre- and post-smoothing sweeps of the mixed V-cycle on the levels
    // above its single-workgroup kernel (0: the XML's effective count) and
    // their Jacobi weight (0: the XML's)
    void set_mixed_sweeps(int pre, int post, double weight = 0.0);
    // n_relax of the single-workgroup kernel in the mixed V-cycle (0: the
    // XML's); when set, its sweeps use the mixed Jacobi weight as well
    void set_mixed_tail(int n_relax);
    // The coarse solve of the single-workgroup kernel when the V-cycle is a
    // preconditioner (vcycle_mixed, vcycle_eager): coarse_tol as mg_tail's
    // (0 the XML's absolute tolerance, > 0 a reduction relative to its
    // start, < 0 a fixed number of sweeps), max_iter > 0 its sweep cap (0:
    // kPrecondCoarseSweeps for a fixed count, else the XML's).  A fixed
    // count keeps the preconditioner one linear operator; with the XML's
    // absolute tolerance it does fewer sweeps as GCR's residual shrinks,
    // which can stall GCR.
    static constexpr int kPrecondCoarseSweeps = 5;
    void set_precond_coarse(double coarse_tol, int max_iter);
    void vcycle_mixed(const double *r);
    float *mixed_result() const { return m_err1f.empty() ? nullptr : m_err1f[0]; }
    // The coarse part of the mixed V-cycle (the levels of the single-workgroup
    // kernel) replaced by the exact solve on its first level, x = K^-1 dsign b
    // with K the level's operator and linear fill, applied as a dense matrix
    // (k::dense_apply) when the level has at most kMaxDirectCells fluid cells,
    // uses the fused fill, holds no obstacle cells and K is not singular
    // beyond a constant null space.
    // On by default; set before the first mixed V-cycle (prepare_mixed
    // assembles it; a GcrPressure built afterwards captures the choice).
    static constexpr int kMaxDirectCells = 1536;
    void set_direct_coarse(bool on);
    bool direct_coarse() const { return m_direct != nullptr; }

    // graphs need the fused schedule with the single-workgroup kernel (no
    // host round trip inside a V-cycle); while a kernel timer is armed the
    // V-cycle runs eagerly so that every launch can be bracketed
    bool graphable() const {
        return m_fused && m_use_graphs && m_tail_first < m_s.n_level && m_domain.timer_tag().empty();
    }

  private:
    struct CycleGraph {
        hipGraphExec_t exec = nullptr;
        double *out = nullptr;
        const double *b = nullptr;
    };

    void launch_graph(int which, Field &out, const Field &b);
    void drop_graphs();
    void residual_async(const Field &out, const Field &b);
    void vcycle_plain(Field &out, const Field &b);
    void vcycle_fused(Field &out, const Field &b);
    void smooth_fused(int level, const double *b);
    void smooth_fast(int level, const double *b);
    bool fast_level(int level) const;
    // Restrict from `level` reads the fine ghost layer (an axis of one cell)
    bool restrict_reads_ghost(int level) const;
    void bc(double *f, int level);
    void build_tail();
    void smooth(Field &out, Field &tmp, const Field &b, int level);
    void solve(Field &out, Field &tmp, const Field &b, int level);
    void smooth_jacobi(Field &out, Field &tmp, const Field &b, int level);
    void solve_jacobi(Field &out, Field &tmp, const Field &b, int level);
    void smooth_cgs(Field &out, const Field &b, int level);
    void solve_cgs(Field &out, const Field &b, int level);
    void coeffs(int level, double *ax, double *ay, double *az, double *beta) const;
    double level_residual(const Field &out, const Field &b, int level);

    MultigridSettings m_s;
    bool m_cgs;
    static constexpr double kDsign = -1;  // VCycleMG.h:53
    // per-level scratch, VCycleMG.cpp:50-98 (residuum1[0] aliases b)
    std::vector<std::unique_ptr<Field>> m_res0, m_res1, m_err0, m_err1, m_tmp;
    int m_cycles = 0;
    bool m_fused = true;
    long long m_tail_cells = 4096;
    int m_tail_first = 0;          // first level of the single-workgroup kernel; n_level: none
    TailPlan *m_plan_dev = nullptr;     // [0] the reference's V-cycle, [1] with the coarsest level skipped
    int m_tail_lds_bytes[2] = {0, 0};
    // FP32 level arrays of the mixed V-cycle (levels above the tail)
    std::vector<float *> m_err1f, m_tmpf, m_scrf, m_res0f, m_res1f;
    int m_mixed_pre = 0, m_mixed_post = 0;  // sweeps of the mixed V-cycle (0: the XML's)
    double m_mixed_weight = 0.0;            // their Jacobi weight (0: the XML's)
    int m_mixed_tail = 0;                   // n_relax of the tail kernel in the mixed V-cycle (0: the XML's)
    double m_pc_tol = -1.0;                 // preconditioner coarse solve: stopping test
    int m_pc_iter = 0;                      // ... and sweep cap (0: default)
    bool m_precond_call = false;            // inside vcycle_eager
    int precond_coarse_iter() const {
        return m_pc_iter > 0 ? m_pc_iter : (m_pc_tol < 0.0 ? kPrecondCoarseSweeps : 0);
    }
    void alloc_mixed();
    void free_mixed();
    void smooth_fast_f(int level, const double *b64, const float *b32, int sweeps, bool from_zero, bool first_done);
    int mixed_sweep_count(int sweeps) const;                  // effective sweeps for a count (0: the XML's)
    float *first_sweep_target(int level, int sweeps) const;   // buffer the first of them writes
    bool build_galerkin();
    bool build_direct();
    void free_direct();
    bool m_direct_enabled = true;
    bool m_direct_tried = false;
    double *m_direct = nullptr;           // n x n row-major, dsign K^-1
    uint32_t *m_direct_cells = nullptr;   // the level's fluid cells (row/column order)
    int m_direct_n = 0;
    bool m_use_graphs = true;
    CycleGraph m_graph[2];  // [0]: V-cycle, [1]: V-cycle + residual
};

}  // namespace artss
