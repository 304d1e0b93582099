// Solver drivers: FieldController (field/FieldController.cpp), NSSolver
// (solver/NSSolver.cpp:50-124), NSTempTurbSolver (solver/NSTempTurbSolver.cpp:94-248),
// NSTempSolver (solver/NSTempSolver.cpp:69-198), NSTurbSolver, DiffusionSolver,
// AdvectionDiffusionSolver, DiffusionTurbSolver and AdvectionSolver (their do_step),
// SolverController setup and source updates (solver/SolverController.cpp:26-311)
// and one TimeIntegration iteration (TimeIntegration.cpp:86-149).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "operators.h"
#include "krylov.h"
#include "unet.h"
#include "vcycle.h"

namespace artss {

struct FieldController {
    // concentration: also the concentration fields C, C0, C_tmp, its source
    // S_C and eddy diffusivity gamma_t (the concentration solvers only)
    FieldController(Domain &d, bool concentration = false);
    Field u, v, w, u0, v0, w0, u_tmp, v_tmp, w_tmp;
    Field nu_t, kappa_t;
    Field p, rhs;
    Field T, T0, T_tmp, T_ambient;
    Field f_x, f_y, f_z, S_T;
    std::unique_ptr<Field> C, C0, C_tmp, S_C, gamma_t;

    // FieldController.cpp:151-186
    void couple_vector();
    void couple_scalar();
    void couple_concentration();
    // FieldController.cpp:119-135
    void update_data();
    Field *by_name(const std::string &name);
    // every field, in a fixed order, with its name (checkpoints)
    std::vector<std::pair<const char *, Field *>> all();
};

// Counters of the last time step: diffusion iterations per component and
// V-cycles of the pressure solve.
struct StepCounters {
    int diffusion[4] = {0, 0, 0, 0};  // u, v, w, T
    int vcycles = 0;
    double residual = -1.0;           // final pressure residual when measured
    double initial_residual = -1.0;   // pressure residual at the start of the iteration (GCR methods)
    bool converged = true;            // a residual-target solve met its target (always true otherwise)
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
    // methods 1, 2: sweep cap of the preconditioner's coarse solve (0: five
    // sweeps with a fixed count, else the XML's max_iter)
    int coarse_max_iter = 0;
    int mixed_pre = 0, mixed_post = 0;  // method 2: smoothing sweeps above the coarse levels (0: the XML's)
    double mixed_weight = 0.0;          // method 2: their Jacobi weight (0: the XML's)
    bool coarse_direct = true;          // method 2: dense exact coarse solve where it applies (VCycleMG)
    int mixed_tail = 0;                 // method 2: n_relax of the coarse levels' kernel (0: the XML's)
    // methods 1, 2: stopping test of the preconditioner's coarse solve: 0
    // the XML's absolute tolerance, > 0 a reduction relative to its starting
    // residual, < 0 (default) a fixed number of sweeps (coarse_max_iter)
    double coarse_tol = -1.0;
};

// Where the residual-target GCR pressure iteration starts (methods 1, 2),
// from p_n (the previous step's solution) and p_prev (the one before):
//   0  p_n, the start the reference's iteration uses
//   1  p_n + d with d = p_n - p_prev (linear extrapolation in time)
//   2  p_n + d + s * net(c r / s, d / s), the warm-start U-Net's correction
//      to the extrapolation, r = rhs - lap(p_n + d) the residual of the
//      extrapolated start, s = rms(d) over the fluid cells,
//      c = 1 / (2/dx^2 + 2/dy^2 + 2/dz^2)
//   3  p_n + sum_s alpha_s d_s over the last `depth` increments d_s of the
//      solution (d_1 = d), the combination closest to the solution in the
//      energy norm (Galerkin projection onto the previous steps' increments;
//      alpha = (1, 0, ...) is the extrapolation, so its energy-norm error is
//      never larger)
// The iteration then runs to the same residual target from any start.
// `record` keeps the last step's network input and target, (p_{n+1} - p_n -
// d) / s, for training.
struct WarmStart {
    int mode = 0;
    UNet *net = nullptr;  // mode 2, not owned
    bool record = false;
    int depth = 4;  // mode 3: increments kept (1 .. kk::kProjMax)
};

class Solver {
  public:
    Solver(std::unique_ptr<Domain> domain);
    ~Solver();

    Domain &domain() { return *m_domain; }
    FieldController &fields() { return *m_fields; }

    // SolverController::solver_do_step + update_sources + FieldController::update_data
    void step(double t);

    // ---- step graph (no reference counterpart).  After one eager step that
    // records how many iterations each residual-target loop took, the whole
    // step is captured in one HIP graph, with that count + margin iterations
    // per loop behind the loop's device gate, and replayed: the host launches
    // it and reads one outcome per loop.  The graph first copies the fields
    // aside; a replay in which a loop would have needed more than its budget
    // (or restarted, or stalled) is rolled back from that copy and the step
    // runs eagerly, after which the graph is captured again.  The fields are
    // those of the eager step bit for bit.  It needs the residual-target GCR
    // pressure solve (methods 1, 2), warm start 0 or 1 without recording, and
    // loops with a capturable form (colored Gauss-Seidel with an idempotent
    // fill); other configurations step eagerly and graph_stats().note says
    // why.  While a kernel timer is armed steps run eagerly.
    void set_step_graph(bool on, int margin);
    struct GraphStats {
        long long replays = 0, rollbacks = 0, captures = 0;
        bool active = false;  // a captured graph is in use
        std::string note;     // why the step is not captured, if it is not
    };
    const GraphStats &graph_stats() const { return m_graph_stats; }
    // a change of the pressure solver's launch schedule makes a captured step stale
    void invalidate_step_graph() { drop_step_graph(); }
    const StepCounters &counters() const { return m_counters; }
    // steps whose GCR pressure solve stalled and was finished by V-cycles
    long long fallbacks() const { return m_fallbacks; }
    void set_pressure_mode(const PressureMode &m) {
        m_pmode = m;
        drop_step_graph();
    }
    const PressureMode &pressure_mode() const { return m_pmode; }
    VCycleMG &pressure_solver() { return *m_pres; }

    // Analysis::save_variables_in_file (analysis/Analysis.cpp:360-408): for
    // every field type with domain boundary data, `<dir>/<name>.dat` with one
    // value per cell in the default ostream format, and, with obstacles,
    // `<dir>/<name>_obstacle.dat` with "index;value" per obstacle cell.
    // Returns the number of files written.
    int write_dat(const std::string &dir);

    // the start of the GCR pressure iteration; throws config_error when
    // mode 2 or `record` is asked for on an inner box whose sides are not
    // multiples of 8 (the network halves it three times)
    void set_warm_start(const WarmStart &w);
    // copies the last recorded sample into device buffers x [V][2] and t [V]
    // (V = inner box cells, x fastest); false when none is recorded yet
    bool warm_sample(float *x, float *t);

    // ---- checkpoint / rollback (the reference's FieldIO and DataAssimilation
    // rollback, dataAssimilation/FieldIO.cpp:44-100, DataAssimilation.cpp:56-161)
    // The state between two steps is every FieldController field plus the
    // warm start's previous solution; restoring it continues a run bit for
    // bit as if it had not been interrupted.
    // in HBM: returns a slot id; restore(slot) rolls the solver back to it
    int snapshot(double t);
    double restore(int slot);
    void drop_snapshot(int slot);
    // on disk: one file per call, a header (magic, version, cell counts,
    // time, time step, field names) followed by the raw FP64 arrays
    void save(const std::string &path, double t);
    double load(const std::string &path);

  private:
    struct Snapshot {
        double t = 0.0;
        std::vector<double *> data;  // device copies, in FieldController::all() order, then p_prev
        int warm_history = 0;
    };
    std::vector<std::unique_ptr<Snapshot>> m_snapshots;
    void warm_begin();
    void warm_end();
    void step_eager(double t);
    bool step_replay(double t);
    void step_capture();
    void drop_step_graph();
    bool m_graph_on = false;
    int m_graph_margin = 2;
    hipGraphExec_t m_step_exec = nullptr;
    LoopLog *m_loop_log = nullptr;       // device
    LoopLog *m_loop_log_host = nullptr;  // pinned
    std::vector<int> m_loop_kind;        // StepCapture kinds of the captured loops
    // the fields (and the warm start's previous solution) and their copies,
    // as device arrays of pointers for one copy launch each way
    std::vector<double *> m_rollback;
    double **m_state_dev = nullptr, **m_rollback_dev = nullptr;
    int m_state_count = 0;
    GraphStats m_graph_stats;
    void do_step_ns(double t);
    void do_step_nstempturb(double t);
    void do_step_nstemp(double t);
    void do_step_nsturb(double t);
    void do_step_concentration();
    void do_step_transport();
    void pressure_step(double t);
    void vcycles_to_tolerance(int &cycles, double &r);
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
    int m_gcr_coarse = 0, m_gcr_pre = 0, m_gcr_post = 0;
    double m_gcr_weight = 0.0;
    bool m_gcr_direct = true, m_gcr_built = false;
    long long m_fallbacks = 0;
    double m_gcr_coarse_tol = 0.0;
    int m_gcr_tail = 0;
    std::unique_ptr<ExplicitEulerSource> m_src_vel, m_src_temp;
    std::unique_ptr<GaussFunction> m_gauss;
    std::unique_ptr<BuoyancyMMS> m_mms;
    // 0 NSSolver, 1 NSTempTurbSolver, 2 PressureSolver, 3 NSTempSolver,
    // 4 NSTurbSolver, 5 DiffusionSolver, 6 AdvectionDiffusionSolver,
    // 7 DiffusionTurbSolver, 8 AdvectionSolver, 9 NSTempConSolver,
    // 10 NSTempTurbConSolver
    int m_kind = 0;
    // the concentration solvers (9, 10): advection, diffusion (the velocity's
    // settings, NSTempTurbConSolver.cpp:35), source and its function
    std::unique_ptr<IAdvection> m_adv_con;
    std::unique_ptr<IDiffusion> m_dif_con;
    std::unique_ptr<ExplicitEulerSource> m_src_con;
    std::unique_ptr<GaussFunction> m_gauss_con;
    std::unique_ptr<Field> m_cube_con;  // Cube source: its constant field
    bool m_add_con_source = false;
    std::unique_ptr<Field> m_lin[3];  // AdvectionSolver's constant velocity
    bool m_add_source = false, m_add_temp_source = false;
    StepCounters m_counters;
    PressureMode m_pmode;
    WarmStart m_warm;
    int m_warm_history = 0;  // solves since the warm start was set
    bool m_warm_sample = false;
    double *m_p_prev = nullptr, *m_p_start = nullptr;
    double *m_warm_partial = nullptr, *m_warm_sum = nullptr;
    // mode 3: increments and their images in a ring of slots, the Gram matrix
    // of the images by slot, the coefficients, partial sums; the slot of the
    // newest increment and the slots that hold one
    kk::ProjRing m_proj{};
    double *m_proj_G = nullptr, *m_proj_alpha = nullptr, *m_proj_partial = nullptr;
    int m_proj_fresh = -1;
    unsigned m_proj_valid = 0;
    void proj_reset() {
        m_proj_fresh = -1;
        m_proj_valid = 0;
    }
    float *m_warm_x = nullptr, *m_warm_y = nullptr, *m_warm_t = nullptr;
};


}  // namespace artss
