// Geometric multigrid V-cycle pressure solver (pressure/VCycleMG.cpp).
#pragma once

#include <memory>
#include <vector>

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

    int last_cycles() const { return m_cycles; }

  private:
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
};

}  // namespace artss
