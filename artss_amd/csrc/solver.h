// Solver drivers: FieldController (field/FieldController.cpp), NSSolver
// (solver/NSSolver.cpp:50-124), NSTempTurbSolver (solver/NSTempTurbSolver.cpp:94-248),
// SolverController setup and source updates (solver/SolverController.cpp:26-311)
// and one TimeIntegration iteration (TimeIntegration.cpp:86-149).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "operators.h"
#include "krylov.h"
#include "vcycle.h"

namespace artss {

struct FieldController {
    explicit FieldController(Domain &d);
    Field u, v, w, u0, v0, w0, u_tmp, v_tmp, w_tmp;
    Field nu_t, kappa_t;
    Field p, rhs;
    Field T, T0, T_tmp, T_ambient;
    Field f_x, f_y, f_z, S_T;

    // FieldController.cpp:151-186
    void couple_vector();
    void couple_scalar();
    // FieldController.cpp:119-135 (concentration fields are not part of this path)
    void update_data();
    Field *by_name(const std::string &name);
};

// Counters of the last time step: diffusion iterations per component and
// V-cycles of the pressure solve.
struct StepCounters {
    int diffusion[4] = {0, 0, 0, 0};  // u, v, w, T
    int vcycles = 0;
    double residual = -1.0;           // final pressure residual when measured
};

// How the pressure equation is solved each step.
//   reference: VCycleMG::pressure (residual loop in step 1, n_cycle afterwards)
//   residual:  V-cycles until ||rhs - lap p|| <= tolerance every step
struct PressureMode {
    bool to_tolerance = false;
    double tolerance = 1e-8;
    int max_cycles = 200;
    // with to_tolerance: 0 V-cycles on their own (the reference's
    // iteration), 1 V-cycles as the preconditioner of GCR (krylov.h),
    // 2 as 1 with the V-cycle's upper levels in FP32
    int method = 0;
    int coarse_max_iter = 0;  // method 2: cap of the preconditioner's coarse solve (0: the XML's)
};

class Solver {
  public:
    Solver(std::unique_ptr<Domain> domain);
    ~Solver();

    Domain &domain() { return *m_domain; }
    FieldController &fields() { return *m_fields; }

    // SolverController::solver_do_step + update_sources + FieldController::update_data
    void step(double t);
    const StepCounters &counters() const { return m_counters; }
    void set_pressure_mode(const PressureMode &m) { m_pmode = m; }
    const PressureMode &pressure_mode() const { return m_pmode; }
    VCycleMG &pressure_solver() { return *m_pres; }

    // Analysis::save_variables_in_file (analysis/Analysis.cpp:360-408): for
    // every field type with domain boundary data, `<dir>/<name>.dat` with one
    // value per cell in the default ostream format, and, with obstacles,
    // `<dir>/<name>_obstacle.dat` with "index;value" per obstacle cell.
    // Returns the number of files written.
    int write_dat(const std::string &dir);

  private:
    void do_step_ns(double t);
    void do_step_nstempturb(double t);
    void pressure_step(double t);
    void update_sources(double t);
    void set_up_fields();

    std::unique_ptr<Domain> m_domain;
    std::unique_ptr<FieldController> m_fields;
    std::unique_ptr<IAdvection> m_adv;
    std::unique_ptr<IDiffusion> m_dif;
    std::unique_ptr<IDiffusion> m_dif_temp;
    std::unique_ptr<ITurbulence> m_turb;
    std::unique_ptr<VCycleMG> m_pres;
    std::unique_ptr<GcrPressure> m_gcr;
    bool m_gcr_mixed = false;
    int m_gcr_coarse = 0;
    std::unique_ptr<ExplicitEulerSource> m_src_vel, m_src_temp;
    std::unique_ptr<GaussFunction> m_gauss;
    int m_kind = 0;  // 0 NSSolver, 1 NSTempTurbSolver, 2 PressureSolver
    bool m_add_source = false, m_add_temp_source = false;
    StepCounters m_counters;
    PressureMode m_pmode;
};

// initial conditions on the inner list (Functions.cpp) for the supported usr_fct values
std::vector<double> initial_condition(const Domain &d, const std::string &what);

}  // namespace artss
