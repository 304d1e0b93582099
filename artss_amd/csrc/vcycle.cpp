#include "vcycle.h"

#include <cmath>

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
}

VCycleMG::~VCycleMG() = default;

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

// VCycleMG::VCycleMultigrid :215-288.  residuum1[0] is the right-hand side
// itself (the reference copies b there and never writes it), and the copy
// of error0[0] into out at :254 is dead (out is overwritten at :282).
void VCycleMG::vcycle(Field &out, const Field &b) {
    Domain &d = m_domain;
    const int L = m_s.n_level;
    if (L == 0) {
        solve(out, *m_tmp[0], b, 0);
        return;
    }
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
