#include "solver.h"

#include <mutex>
#include <shared_mutex>

#include <fstream>
#include <utility>

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "functions.h"
#include "krylov_kernels.h"

namespace artss {

FieldController::FieldController(Domain &d, bool concentration)
    : u(d, U), v(d, V), w(d, W), u0(d, U), v0(d, V), w0(d, W), u_tmp(d, U), v_tmp(d, V), w_tmp(d, W),
      nu_t(d, NU), kappa_t(d, artss::T), p(d, P), rhs(d, P), T(d, artss::T), T0(d, artss::T), T_tmp(d, artss::T),
      T_ambient(d, artss::T), f_x(d, U), f_y(d, V), f_z(d, W), S_T(d, artss::T) {
    if (concentration) {
        // FieldController.cpp:44-53 (concentration, its copies and source are
        // field type RHO; gamma_t, like kappa_t, carries T's boundary data)
        C = std::make_unique<Field>(d, RHO);
        C0 = std::make_unique<Field>(d, RHO);
        C_tmp = std::make_unique<Field>(d, RHO);
        S_C = std::make_unique<Field>(d, RHO);
        gamma_t = std::make_unique<Field>(d, RHO);
    }
}

void FieldController::couple_concentration() {
    k::CopyFan f{{C->data}, {C0->data}, {C_tmp->data}, 1};
    k::copy_fan(C->domain().stream(), f, C->get_size());
}

// FieldController.cpp couple_vector / couple_scalar / update_data: the copies
// into the previous-step and temporary fields, one pass per call
void FieldController::couple_vector() {
    k::CopyFan f{{u.data, v.data, w.data}, {u0.data, v0.data, w0.data}, {u_tmp.data, v_tmp.data, w_tmp.data}, 3};
    k::copy_fan(u.domain().stream(), f, u.get_size());
}

void FieldController::couple_scalar() {
    k::CopyFan f{{T.data}, {T0.data}, {T_tmp.data}, 1};
    k::copy_fan(T.domain().stream(), f, T.get_size());
}

void FieldController::update_data() {
    k::CopyFan f{{u.data, v.data, w.data, T.data},
                 {u0.data, v0.data, w0.data, T0.data},
                 {u_tmp.data, v_tmp.data, w_tmp.data, T_tmp.data},
                 4};
    k::copy_fan(u.domain().stream(), f, u.get_size());
    if (C) couple_concentration();
}

std::vector<std::pair<const char *, Field *>> FieldController::all() {
    std::vector<std::pair<const char *, Field *>> out = {
        {"u", &u},         {"v", &v},         {"w", &w},         {"u0", &u0},   {"v0", &v0},
        {"w0", &w0},       {"u_tmp", &u_tmp}, {"v_tmp", &v_tmp}, {"w_tmp", &w_tmp},
        {"nu_t", &nu_t},   {"kappa_t", &kappa_t}, {"p", &p},     {"rhs", &rhs}, {"T", &T},
        {"T0", &T0},       {"T_tmp", &T_tmp}, {"T_ambient", &T_ambient}, {"f_x", &f_x},
        {"f_y", &f_y},     {"f_z", &f_z},     {"S_T", &S_T}};
    if (C) {
        out.push_back({"C", C.get()});
        out.push_back({"C0", C0.get()});
        out.push_back({"C_tmp", C_tmp.get()});
        out.push_back({"S_C", S_C.get()});
        out.push_back({"gamma_t", gamma_t.get()});
    }
    return out;
}

Field *FieldController::by_name(const std::string &n) {
    // Analysis.cpp:368: the RHO field is the concentration
    if (C && (n == "C" || n == "rho")) return C.get();
    if (C && n == "C0") return C0.get();
    if (C && n == "C_tmp") return C_tmp.get();
    if (C && n == "S_C") return S_C.get();
    if (C && n == "gamma_t") return gamma_t.get();
    if (n == "u") return &u;
    if (n == "v") return &v;
    if (n == "w") return &w;
    if (n == "u0") return &u0;
    if (n == "v0") return &v0;
    if (n == "w0") return &w0;
    if (n == "u_tmp") return &u_tmp;
    if (n == "v_tmp") return &v_tmp;
    if (n == "w_tmp") return &w_tmp;
    if (n == "p") return &p;
    if (n == "rhs") return &rhs;
    if (n == "T") return &T;
    if (n == "T0") return &T0;
    if (n == "T_tmp") return &T_tmp;
    if (n == "T_ambient") return &T_ambient;
    if (n == "nu_t") return &nu_t;
    if (n == "kappa_t") return &kappa_t;
    if (n == "f_x") return &f_x;
    if (n == "f_y") return &f_y;
    if (n == "f_z") return &f_z;
    if (n == "S_T") return &S_T;
    return nullptr;
}

namespace {
void upload(Domain &d, double *dst, const double *src, size_t n) {
    ARTSS_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, d.stream()));
}
}  // namespace

Solver::Solver(std::unique_ptr<Domain> domain) : m_domain(std::move(domain)) {
    Domain &d = *m_domain;
    const Settings &s = d.settings();
    static const char *kinds[] = {"NSSolver",        "NSTempTurbSolver",         "PressureSolver",
                                  "NSTempSolver",    "NSTurbSolver",             "DiffusionSolver",
                                  "AdvectionDiffusionSolver", "DiffusionTurbSolver", "AdvectionSolver",
                                  "NSTempConSolver", "NSTempTurbConSolver"};
    m_kind = -1;
    for (int k = 0; k < 11; k++)
        if (s.solver == kinds[k]) m_kind = k;
    if (m_kind < 0)
        throw config_error("solver '" + s.solver + "' is not supported by this library (NSSolver, NSTurbSolver, NSTempSolver, "
                           "NSTempTurbSolver, NSTempConSolver, NSTempTurbConSolver, PressureSolver, DiffusionSolver, "
                           "DiffusionTurbSolver, AdvectionDiffusionSolver, AdvectionSolver)");
    const bool con = m_kind == 9 || m_kind == 10;
    const bool temp_turb = m_kind == 1 || m_kind == 10;  // the NSTempTurb step
    const bool temp_lam = m_kind == 3 || m_kind == 9;    // the NSTemp step
    if (m_kind == 8 && s.ic != "GaussBubble")
        throw config_error("AdvectionSolver takes its velocity from the GaussBubble initial condition "
                           "(SolverController.cpp:54-56)");
    if (s.has_temperature && s.temp_random) throw config_error("random temperature source is not supported");
    if (s.has_temperature && s.dissipation && m_kind != 3 && !con)
        throw config_error("dissipation is part of NSTempSolver and the concentration solvers only "
                           "(NSTempSolver.cpp:181-185, NSTempConSolver.cpp, NSTempTurbConSolver.cpp)");

    m_fields = std::make_unique<FieldController>(d, con);
    // the solver constructors (SolverSelection::set_*): NSSolver.cpp:17-36,
    // NSTurbSolver.cpp, NSTempSolver.cpp:16-49, NSTempTurbSolver.cpp,
    // DiffusionSolver.cpp, AdvectionDiffusionSolver.cpp, DiffusionTurbSolver.cpp
    const bool velocity_adv = m_kind == 0 || temp_turb || temp_lam || m_kind == 4 || m_kind == 6 || m_kind == 8;
    const bool velocity_dif = m_kind != 2 && m_kind != 8;
    const bool ns = m_kind == 0 || temp_turb || temp_lam || m_kind == 4;
    if (velocity_adv) m_adv = make_advection(d, s.advection_type);
    if (velocity_dif) m_dif = make_diffusion(d, s.diffusion);
    if (ns) {
        m_src_vel = std::make_unique<ExplicitEulerSource>(d, s.force_dir);
        m_add_source = s.force_fct != "Zero";
    }
    if (ns || m_kind == 2) m_pres = std::make_unique<VCycleMG>(d, s.pressure);
    if (temp_turb || m_kind == 4 || m_kind == 7) m_turb = make_turbulence(d, s);
    if (m_kind == 8) {
        // AdvectionSolver.cpp:16-34: constant transport velocity from the GaussBubble parameters
        const char *keys[3] = {"u_lin", "v_lin", "w_lin"};
        for (int a = 0; a < 3; a++) {
            auto it = s.ic_values.find(keys[a]);
            if (it == s.ic_values.end()) throw config_error(std::string("GaussBubble needs <") + keys[a] + ">");
            m_lin[a] = std::make_unique<Field>(d, U + a, std::strtod(it->second.c_str(), nullptr));
        }
    }
    if (temp_turb || temp_lam) {
        if (s.temp_advection_type != "SemiLagrangian") throw config_error("temperature advection must be SemiLagrangian");
        m_dif_temp = make_diffusion(d, s.diffusion);  // both solvers use the velocity settings (NSTempSolver.cpp:33)
        m_src_temp = std::make_unique<ExplicitEulerSource>(d, s.temp_dir);
        if (s.temp_fct == "Gauss") m_gauss = std::make_unique<GaussFunction>(d, s.gauss);
        if (s.temp_fct == "Buoyancy") m_mms = std::make_unique<BuoyancyMMS>(d);  // SolverSelection.cpp:148-149
        m_add_temp_source = s.temp_fct != "Zero";
    }
    if (con) {
        // NSTempTurbConSolver.cpp:15-58, NSTempConSolver.cpp: the reference
        // hands the concentration's advection to adv_temp a second time and
        // leaves adv_con unset, so its step calls through an uninitialised
        // pointer; here the concentration is advected by the method its own
        // section names.
        const ConcentrationSettings &c = s.concentration;
        m_adv_con = make_advection(d, c.advection_type);
        m_dif_con = make_diffusion(d, s.diffusion);
        m_src_con = std::make_unique<ExplicitEulerSource>(d, c.dir);
        if (c.con_fct == "Gauss") {
            m_gauss_con = std::make_unique<GaussFunction>(d, c.gauss);
        } else {
            m_cube_con = std::make_unique<Field>(d, RHO);
            const std::vector<double> host = d.cube_source(c.cube);
            upload(d, m_cube_con->data, host.data(), host.size());
            ARTSS_HIP(hipStreamSynchronize(d.stream()));
        }
        m_add_con_source = true;  // con_fct is Gauss or Cube (Settings.cpp:880-887)
    }
    set_up_fields();
}

Solver::~Solver() {
    if (m_domain) (void)hipStreamSynchronize(m_domain->stream());
    if (m_step_exec) (void)hipGraphExecDestroy(m_step_exec);
    for (double *q : m_rollback) (void)hipFree(q);
    for (void *q : {static_cast<void *>(m_loop_log), static_cast<void *>(m_state_dev),
                    static_cast<void *>(m_rollback_dev)})
        if (q) (void)hipFree(q);
    if (m_loop_log_host) (void)hipHostFree(m_loop_log_host);
    for (auto &snap : m_snapshots)
        if (snap)
            for (double *q : snap->data) (void)hipFree(q);
    for (void *q : {static_cast<void *>(m_p_prev), static_cast<void *>(m_p_start),
                    static_cast<void *>(m_warm_partial), static_cast<void *>(m_warm_sum),
                    static_cast<void *>(m_warm_x), static_cast<void *>(m_warm_y), static_cast<void *>(m_warm_t),
                    static_cast<void *>(m_proj_G), static_cast<void *>(m_proj_alpha),
                    static_cast<void *>(m_proj_partial)})
        if (q) (void)hipFree(q);
    for (int q = 0; q < kk::kProjMax; q++) {
        if (m_proj.d[q]) (void)hipFree(m_proj.d[q]);
        if (m_proj.w[q]) (void)hipFree(m_proj.w[q]);
    }
}

void Solver::set_warm_start(const WarmStart &w) {
    drop_step_graph();
    if (w.mode < 0 || w.mode > 3) throw config_error("warm start mode must be 0, 1, 2 or 3");
    if (w.mode == 3 && (w.depth < 1 || w.depth > kk::kProjMax))
        throw config_error("the projection start keeps 1 to " + std::to_string(kk::kProjMax) + " increments");
    if (w.mode == 2 && w.net == nullptr) throw config_error("the learned warm start needs a network");
    const LevelDesc &L = m_domain->level(0);
    const bool net = w.mode == 2 || w.record;
    if (net && (L.n[0] % 8 || L.n[1] % 8 || L.n[2] % 8))
        throw config_error("the warm-start network needs inner box sides that are multiples of 8");
    const size_t n = static_cast<size_t>(L.size);
    const size_t v = static_cast<size_t>(L.n[0]) * L.n[1] * L.n[2];
    if (!m_p_prev) {
        ARTSS_HIP(hipMalloc(&m_p_prev, sizeof(double) * n));
        ARTSS_HIP(hipMalloc(&m_p_start, sizeof(double) * n));
        ARTSS_HIP(hipMalloc(&m_warm_partial, sizeof(double) * static_cast<size_t>(kk::blocks_inner(L))));
        ARTSS_HIP(hipMalloc(&m_warm_sum, sizeof(double)));
    }
    if (w.mode == 3) {
        for (int q = 0; q < w.depth; q++)
            if (!m_proj.d[q]) {
                ARTSS_HIP(hipMalloc(&m_proj.d[q], sizeof(double) * n));
                ARTSS_HIP(hipMalloc(&m_proj.w[q], sizeof(double) * n));
            }
        if (!m_proj_G) {
            ARTSS_HIP(hipMalloc(&m_proj_G, sizeof(double) * kk::kProjMax * kk::kProjMax));
            ARTSS_HIP(hipMalloc(&m_proj_alpha, sizeof(double) * kk::kProjMax));
            ARTSS_HIP(hipMalloc(&m_proj_partial,
                                sizeof(double) * 2 * kk::kProjMax * static_cast<size_t>(kk::blocks_inner(L))));
        }
        // the increments restart from the previous solution
        proj_reset();
    }
    if (net && !m_warm_x) {
        ARTSS_HIP(hipMalloc(&m_warm_x, sizeof(float) * 2 * v));
        ARTSS_HIP(hipMalloc(&m_warm_y, sizeof(float) * v));
        ARTSS_HIP(hipMalloc(&m_warm_t, sizeof(float) * v));
    }
    // the previous solution stays valid across a change of mode when it
    // was being kept
    const bool kept = m_warm.mode != 0 || m_warm.record;
    m_warm = w;
    if (!kept) m_warm_history = 0;
    if (!w.record) m_warm_sample = false;
}

bool Solver::warm_sample(float *x, float *t) {
    if (!m_warm_sample) return false;
    const LevelDesc &L = m_domain->level(0);
    const size_t v = static_cast<size_t>(L.n[0]) * L.n[1] * L.n[2];
    hipStream_t s = m_domain->stream();
    ARTSS_HIP(hipMemcpyAsync(x, m_warm_x, sizeof(float) * 2 * v, hipMemcpyDeviceToDevice, s));
    ARTSS_HIP(hipMemcpyAsync(t, m_warm_t, sizeof(float) * v, hipMemcpyDeviceToDevice, s));
    ARTSS_HIP(hipStreamSynchronize(s));
    return true;
}

// p_n is saved; with a previous solution p is moved to the warm start and
// the network input is kept for recording
void Solver::warm_begin() {
    Domain &d = *m_domain;
    FieldController &f = *m_fields;
    const LevelDesc &L = d.level(0);
    hipStream_t s = d.stream();
    if (d.capture().mode == StepCapture::kCapture && (m_warm.mode >= 2 || m_warm.record))
        throw capture_unsupported("the learned and projection starts and recording");
    k::copy(s, m_p_start, f.p.data, L.size);
    if (m_warm_history == 0) return;
    const double fluid = static_cast<double>(d.inner_cells(0));
    kk::warm_sum(s, L, f.p.data, m_p_prev, m_warm_partial, m_warm_sum);
    const bool net = m_warm.mode == 2 || m_warm.record;
    if (net) {
        const double c = 1.0 / (2.0 / (L.h[0] * L.h[0]) + 2.0 / (L.h[1] * L.h[1]) + 2.0 / (L.h[2] * L.h[2]));
        kk::warm_pack(s, L, f.p.data, f.rhs.data, m_p_prev, m_warm_sum, fluid, c, m_warm_x);
    }
    if (m_warm.mode == 2) {
        d.timed("warm_unet", 0, [&] { m_warm.net->forward(L.n[0], L.n[1], L.n[2], m_warm_x, m_warm_y); });
        kk::warm_apply(s, L, m_warm_sum, fluid, m_warm_y, m_p_prev, f.p.data);
    } else if (m_warm.mode == 1) {
        kk::warm_apply(s, L, m_warm_sum, fluid, nullptr, m_p_prev, f.p.data);
    } else if (m_warm.mode == 3) {
        m_proj_fresh = (m_proj_fresh + 1) % m_warm.depth;
        m_proj_valid |= 1u << m_proj_fresh;
        kk::proj_update(s, L, f.p.data, m_p_prev, f.rhs.data, m_proj, m_proj_fresh, m_proj_valid, m_proj_partial);
        kk::proj_solve(s, m_proj_partial, kk::blocks_inner(L), m_proj_fresh, m_proj_valid, m_proj_G, m_proj_alpha);
        kk::proj_apply(s, L, m_proj, m_proj_fresh, m_proj_valid, m_proj_alpha, f.p.data);
    }
}

// the target of this step (p_{n+1} - p_n - d) / s when recording; p_n
// becomes p_prev of the next solve
void Solver::warm_end() {
    Domain &d = *m_domain;
    const LevelDesc &L = d.level(0);
    if (m_warm.record && m_warm_history > 0) {
        kk::warm_target(d.stream(), L, m_warm_sum, static_cast<double>(d.inner_cells(0)), m_fields->p.data,
                        m_p_start, m_p_prev, m_warm_t);
        m_warm_sample = true;
    }
    if (m_graph_on) {
        // a captured step keeps its buffers: copy instead of exchanging
        k::copy(d.stream(), m_p_prev, m_p_start, L.size);
    } else {
        std::swap(m_p_prev, m_p_start);
    }
    // a captured step counts its solve when a replay succeeds
    if (d.capture().mode != StepCapture::kCapture) m_warm_history++;
}

// SolverController.cpp:26-43, 95-205
void Solver::set_up_fields() {
    Domain &d = *m_domain;
    const Settings &s = d.settings();
    FieldController &f = *m_fields;
    const size_t n = static_cast<size_t>(d.level(0).size);
    // SolverController::set_up_fields (SolverController.cpp:94-196)
    const InitialFields ic = initial_fields(d);
    for (const auto &kv : ic.arrays) {
        Field *fld = f.by_name(kv.first);
        if (fld == nullptr) throw config_error("no field '" + kv.first + "' for the initial condition");
        upload(d, fld->data, kv.second.data(), n);
    }
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
    for (const auto &kv : ic.constants) f.by_name(kv.first)->set_value(kv.second);
    // SolverController::force_source (SolverController.cpp:197-202, 229-257), for the temperature
    // solvers only.  The reference constructs T_ambient and the force fields without a value
    // (FieldController.cpp:41,50-52); they hold zero here, so with use_init_values="Yes" T_ambient
    // stays zero, and the Uniform force of the other solvers adds zero, as the reference's fresh
    // allocations read in practice.
    if (s.has_temperature) {
        if (s.force_fct == "Buoyancy" && !s.use_init_values) f.T_ambient.set_value(s.ambient_temperature);
        if (s.force_fct == "Uniform") {
            // Functions::uniform (Functions.cpp:820-832): the value on the inner cells
            Field *out[3] = {&f.f_x, &f.f_y, &f.f_z};
            const std::vector<uint64_t> inner = d.inner_list(0);
            std::vector<double> host(n);
            for (int a : s.force_dir) {
                std::fill(host.begin(), host.end(), 0.0);
                for (uint64_t idx : inner) host[idx] = s.uniform_force[a];
                upload(d, out[a]->data, host.data(), n);
                ARTSS_HIP(hipStreamSynchronize(d.stream()));
            }
        }
    }
    // FieldController::set_up_boundary, FieldController.cpp:101-112
    d.apply_boundary(f.u.data, U, 0);
    d.apply_boundary(f.v.data, V, 0);
    d.apply_boundary(f.w.data, W, 0);
    d.apply_boundary(f.p.data, P, 0);
    d.apply_boundary(f.T.data, artss::T, 0);
    if (f.C) d.apply_boundary(f.C->data, RHO, 0);
    d.apply_boundary(f.T_ambient.data, artss::T, 0);
    update_sources(0);
    f.update_data();
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
}

// SolverController::update_sources, SolverController.cpp:263-307
void Solver::update_sources(double t) {
    Domain &d = *m_domain;
    const Settings &s = d.settings();
    FieldController &f = *m_fields;
    if (m_src_vel && s.force_fct == "Buoyancy") {
        Field *out[3] = {&f.f_x, &f.f_y, &f.f_z};
        for (int a : s.force_dir) ISource::buoyancy_force(*out[a], f.T, f.T_ambient, s.beta, s.g, false);
    }
    if (m_gauss) m_gauss->update_source(f.S_T, t);
    if (m_mms) m_mms->update_source(f.S_T, t);
    // NSTempTurbConSolver.cpp:358-361: the concentration's source function
    // (GaussFunction.cpp:121-123, Cube.cpp: a copy of its constant field)
    if (m_gauss_con) m_gauss_con->update_source(*f.S_C, t);
    if (m_cube_con) f.S_C->copy_data(*m_cube_con);
}

void Solver::pressure_step(double t) {
    FieldController &f = *m_fields;
    m_counters.converged = true;
    const bool capturing = m_domain->capture().mode == StepCapture::kCapture;
    if (capturing && !(m_pmode.to_tolerance && (m_pmode.method == 1 || m_pmode.method == 2)))
        throw capture_unsupported("a pressure solve other than the residual-target GCR");
    if (!m_pmode.to_tolerance) {
        m_pres->pressure(f.p, f.rhs, t, false);
        m_counters.vcycles = m_pres->last_cycles();
        return;
    }
    if (m_pmode.method == 1 || m_pmode.method == 2) {
        const bool mixed = m_pmode.method == 2;
        if (!m_gcr || m_gcr_mixed != mixed || m_gcr_coarse != m_pmode.coarse_max_iter ||
            m_gcr_pre != m_pmode.mixed_pre || m_gcr_post != m_pmode.mixed_post ||
            m_gcr_weight != m_pmode.mixed_weight || m_gcr_direct != m_pmode.coarse_direct ||
            m_gcr_tail != m_pmode.mixed_tail || m_gcr_coarse_tol != m_pmode.coarse_tol) {
            m_gcr.reset();
            m_pres->set_mixed_sweeps(m_pmode.mixed_pre, m_pmode.mixed_post, m_pmode.mixed_weight);
            m_pres->set_mixed_tail(m_pmode.mixed_tail);
            m_pres->set_precond_coarse(m_pmode.coarse_tol, m_pmode.coarse_max_iter);
            if (m_gcr_direct != m_pmode.coarse_direct || !m_gcr_built) m_pres->set_direct_coarse(m_pmode.coarse_direct);
            m_gcr = std::make_unique<GcrPressure>(*m_domain, *m_pres, mixed);
            m_gcr_mixed = mixed;
            m_gcr_coarse = m_pmode.coarse_max_iter;
            m_gcr_pre = m_pmode.mixed_pre;
            m_gcr_post = m_pmode.mixed_post;
            m_gcr_weight = m_pmode.mixed_weight;
            m_gcr_direct = m_pmode.coarse_direct;
            m_gcr_tail = m_pmode.mixed_tail;
            m_gcr_coarse_tol = m_pmode.coarse_tol;
            m_gcr_built = true;
        }
        const bool warm = m_warm.mode != 0 || m_warm.record;
        if (warm) warm_begin();
        double r = 0.0;
        m_counters.vcycles = m_gcr->solve(f.p, f.rhs, m_pmode.tolerance, m_pmode.max_cycles, &r);
        m_counters.residual = r;
        m_counters.initial_residual = m_gcr->initial_residual();
        m_counters.converged = m_gcr->converged();
        if (!capturing && !m_counters.converged && m_counters.vcycles < m_pmode.max_cycles) {
            // GCR stalled: the reference's V-cycle iteration finishes the
            // solve from where it stopped, within the same cycle budget
            m_fallbacks++;
            int cycles = m_counters.vcycles;
            vcycles_to_tolerance(cycles, r);
            m_counters.vcycles = cycles;
            m_counters.residual = r;
            m_counters.converged = r <= m_pmode.tolerance;
        }
        if (warm) warm_end();
        return;
    }
    int cycles = 0;
    double r = m_pres->residual(f.p, f.rhs);
    vcycles_to_tolerance(cycles, r);
    m_counters.vcycles = cycles;
    m_counters.residual = r;
    m_counters.converged = r <= m_pmode.tolerance;
}

// V-cycles on p until ||rhs - lap p||_2 <= tolerance or the cycle budget
// is spent; `cycles` counts on from its value, `r` holds p's residual
void Solver::vcycles_to_tolerance(int &cycles, double &r) {
    FieldController &f = *m_fields;
    while (r > m_pmode.tolerance && cycles < m_pmode.max_cycles) {
        r = m_pres->vcycle_residual(f.p, f.rhs);
        cycles++;
    }
}

// NSSolver.cpp:50-124
void Solver::do_step_ns(double t) {
    FieldController &f = *m_fields;
    const double nu = m_domain->settings().nu;
    m_adv->advect(f.u, f.u0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.v, f.v0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.w, f.w0, f.u0, f.v0, f.w0, false);
    f.couple_vector();
    if (nu != 0.) {
        m_dif->diffuse(f.u, f.u0, f.u_tmp, nu, false);
        m_counters.diffusion[0] = m_dif->last_iterations();
        m_dif->diffuse(f.v, f.v0, f.v_tmp, nu, false);
        m_counters.diffusion[1] = m_dif->last_iterations();
        m_dif->diffuse(f.w, f.w0, f.w_tmp, nu, false);
        m_counters.diffusion[2] = m_dif->last_iterations();
        if (!m_add_source) f.couple_vector();  // else overwritten below before any read
    }
    if (m_add_source) {
        m_src_vel->add_source(f.u, f.v, f.w, f.f_x, f.f_y, f.f_z, false);
        f.couple_vector();
    }
    m_pres->divergence(f.rhs, f.u_tmp, f.v_tmp, f.w_tmp, false);
    pressure_step(t);
    m_pres->projection(f.u, f.v, f.w, f.u_tmp, f.v_tmp, f.w_tmp, f.p, false);
}

// NSTempTurbSolver.cpp:94-248
void Solver::do_step_nstempturb(double t) {
    Domain &d = *m_domain;
    const Settings &s = d.settings();
    FieldController &f = *m_fields;
    m_adv->advect(f.u, f.u0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.v, f.v0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.w, f.w0, f.u0, f.v0, f.w0, false);
    f.couple_vector();
    m_turb->calc_turbulent_viscosity(f.nu_t, f.u, f.v, f.w, false);
    m_dif->diffuse(f.u, f.u0, f.u_tmp, s.nu, f.nu_t, false);
    m_counters.diffusion[0] = m_dif->last_iterations();
    m_dif->diffuse(f.v, f.v0, f.v_tmp, s.nu, f.nu_t, false);
    m_counters.diffusion[1] = m_dif->last_iterations();
    m_dif->diffuse(f.w, f.w0, f.w_tmp, s.nu, f.nu_t, false);
    m_counters.diffusion[2] = m_dif->last_iterations();
    if (!m_add_source) f.couple_vector();  // else overwritten below before any read
    if (m_add_source) {
        m_src_vel->add_source(f.u, f.v, f.w, f.f_x, f.f_y, f.f_z, false);
        f.couple_vector();
    }
    m_pres->divergence(f.rhs, f.u_tmp, f.v_tmp, f.w_tmp, false);
    pressure_step(t);
    m_pres->projection(f.u, f.v, f.w, f.u_tmp, f.v_tmp, f.w_tmp, f.p, false);
    m_adv->advect(f.T, f.T0, f.u, f.v, f.w, false);
    f.couple_scalar();
    // NSTempTurbConSolver.cpp:149-169 adds the dissipation instead of the
    // source when it is switched on; NSTempTurbSolver has no dissipation
    const bool dissipate = m_kind == 10 && s.dissipation;
    const bool after = dissipate || m_add_temp_source;  // a coupling follows the diffusion's
    if (s.temp_turbulence) {
        const double rPr_T = 1. / s.prandtl;
        k::scale_to(d.stream(), f.kappa_t.data, f.nu_t.data, rPr_T, f.kappa_t.get_size());
        m_dif_temp->diffuse(f.T, f.T0, f.T_tmp, s.kappa, f.kappa_t, false);
        m_counters.diffusion[3] = m_dif_temp->last_iterations();
        if (!after) f.couple_scalar();  // else overwritten below before any read
    } else if (s.kappa != 0.) {
        m_dif_temp->diffuse(f.T, f.T0, f.T_tmp, s.kappa, false);
        m_counters.diffusion[3] = m_dif_temp->last_iterations();
        if (!after) f.couple_scalar();
    }
    if (dissipate) {
        ISource::dissipate(f.T, f.u, f.v, f.w, false);
        f.couple_scalar();
    } else if (m_add_temp_source) {
        m_src_temp->add_source(f.T, f.S_T, false);
        f.couple_scalar();
    }
}

// the concentration step of NSTempConSolver (NSTempConSolver.cpp:204-233) and
// NSTempTurbConSolver (NSTempTurbConSolver.cpp:241-286), after the temperature
void Solver::do_step_concentration() {
    Domain &d = *m_domain;
    const Settings &s = d.settings();
    FieldController &f = *m_fields;
    const ConcentrationSettings &c = s.concentration;
    m_adv_con->advect(*f.C, *f.C0, f.u, f.v, f.w, false);
    f.couple_concentration();
    if (m_kind == 10 && c.turbulence) {
        const double rSc_T = 1. / c.schmidt;
        k::scale_to(d.stream(), f.gamma_t->data, f.nu_t.data, rSc_T, f.gamma_t->get_size());
        m_dif_con->diffuse(*f.C, *f.C0, *f.C_tmp, s.gamma, *f.gamma_t, false);
        f.couple_concentration();
    } else if (s.gamma != 0.) {
        m_dif_con->diffuse(*f.C, *f.C0, *f.C_tmp, s.gamma, false);
        f.couple_concentration();
    }
    // NSTempTurbConSolver adds the source only when the concentration is
    // turbulent (:277-285); NSTempConSolver whenever there is one (:222-230)
    if (m_kind == 10 ? c.turbulence : m_add_con_source) {
        m_src_con->add_source(*f.C, *f.S_C, false);
        f.couple_concentration();
    }
}

// NSTempSolver.cpp:69-198: NSTempTurbSolver without turbulence, laminar
// diffusion, and the optional viscous dissipation source
void Solver::do_step_nstemp(double t) {
    Domain &d = *m_domain;
    const Settings &s = d.settings();
    FieldController &f = *m_fields;
    m_adv->advect(f.u, f.u0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.v, f.v0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.w, f.w0, f.u0, f.v0, f.w0, false);
    f.couple_vector();
    if (s.nu != 0.) {
        m_dif->diffuse(f.u, f.u0, f.u_tmp, s.nu, false);
        m_counters.diffusion[0] = m_dif->last_iterations();
        m_dif->diffuse(f.v, f.v0, f.v_tmp, s.nu, false);
        m_counters.diffusion[1] = m_dif->last_iterations();
        m_dif->diffuse(f.w, f.w0, f.w_tmp, s.nu, false);
        m_counters.diffusion[2] = m_dif->last_iterations();
        if (!m_add_source) f.couple_vector();  // else overwritten below before any read
    }
    if (m_add_source) {
        m_src_vel->add_source(f.u, f.v, f.w, f.f_x, f.f_y, f.f_z, false);
        f.couple_vector();
    }
    m_pres->divergence(f.rhs, f.u_tmp, f.v_tmp, f.w_tmp, false);
    pressure_step(t);
    m_pres->projection(f.u, f.v, f.w, f.u_tmp, f.v_tmp, f.w_tmp, f.p, false);
    m_adv->advect(f.T, f.T0, f.u, f.v, f.w, false);
    f.couple_scalar();
    if (s.kappa != 0.) {
        m_dif_temp->diffuse(f.T, f.T0, f.T_tmp, s.kappa, false);
        m_counters.diffusion[3] = m_dif_temp->last_iterations();
        f.couple_scalar();
    }
    if (s.dissipation) {
        ISource::dissipate(f.T, f.u, f.v, f.w, false);
        f.couple_scalar();
    }
    if (m_add_temp_source) {
        m_src_temp->add_source(f.T, f.S_T, false);
        f.couple_scalar();
    }
}

// NSTurbSolver.cpp do_step: NSSolver with Smagorinsky viscosity in the diffusion
void Solver::do_step_nsturb(double t) {
    const Settings &s = m_domain->settings();
    FieldController &f = *m_fields;
    m_adv->advect(f.u, f.u0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.v, f.v0, f.u0, f.v0, f.w0, false);
    m_adv->advect(f.w, f.w0, f.u0, f.v0, f.w0, false);
    f.couple_vector();
    m_turb->calc_turbulent_viscosity(f.nu_t, f.u, f.v, f.w, false);
    m_dif->diffuse(f.u, f.u0, f.u_tmp, s.nu, f.nu_t, false);
    m_counters.diffusion[0] = m_dif->last_iterations();
    m_dif->diffuse(f.v, f.v0, f.v_tmp, s.nu, f.nu_t, false);
    m_counters.diffusion[1] = m_dif->last_iterations();
    m_dif->diffuse(f.w, f.w0, f.w_tmp, s.nu, f.nu_t, false);
    m_counters.diffusion[2] = m_dif->last_iterations();
    if (!m_add_source) f.couple_vector();  // else overwritten below before any read
    if (m_add_source) {
        m_src_vel->add_source(f.u, f.v, f.w, f.f_x, f.f_y, f.f_z, false);
        f.couple_vector();
    }
    m_pres->divergence(f.rhs, f.u_tmp, f.v_tmp, f.w_tmp, false);
    pressure_step(t);
    m_pres->projection(f.u, f.v, f.w, f.u_tmp, f.v_tmp, f.w_tmp, f.p, false);
}

// DiffusionSolver.cpp, AdvectionDiffusionSolver.cpp, DiffusionTurbSolver.cpp
// and AdvectionSolver.cpp do_step
void Solver::do_step_transport() {
    const Settings &s = m_domain->settings();
    FieldController &f = *m_fields;
    if (m_kind == 8) {
        m_adv->advect(f.u, f.u0, *m_lin[0], *m_lin[1], *m_lin[2], false);
        m_adv->advect(f.v, f.v0, *m_lin[0], *m_lin[1], *m_lin[2], false);
        m_adv->advect(f.w, f.w0, *m_lin[0], *m_lin[1], *m_lin[2], false);
        return;
    }
    if (m_kind == 6) {
        m_adv->advect(f.u, f.u0, f.u0, f.v0, f.w0, false);
        m_adv->advect(f.v, f.v0, f.u0, f.v0, f.w0, false);
        m_adv->advect(f.w, f.w0, f.u0, f.v0, f.w0, false);
        f.couple_vector();
        if (s.nu == 0.) return;
    }
    if (m_kind == 7) m_turb->calc_turbulent_viscosity(f.nu_t, f.u, f.v, f.w, false);
    Field *out[3] = {&f.u, &f.v, &f.w}, *in[3] = {&f.u0, &f.v0, &f.w0}, *b[3] = {&f.u_tmp, &f.v_tmp, &f.w_tmp};
    for (int c = 0; c < 3; c++) {
        if (m_kind == 7) m_dif->diffuse(*out[c], *in[c], *b[c], s.nu, f.nu_t, false);
        else m_dif->diffuse(*out[c], *in[c], *b[c], s.nu, false);
        m_counters.diffusion[c] = m_dif->last_iterations();
    }
}

// A stream capture fails when another thread meanwhile makes a call that
// synchronises with the whole device (a synchronous copy from pageable host
// memory, a free): eager steps, which make such calls when they set things
// up, hold this lock shared, a capture holds it exclusively.  Replays make
// no such call and take no lock.
static std::shared_mutex g_capture_lock;

void Solver::step(double t) {
    Domain &d = *m_domain;
    if (!m_graph_on || !m_graph_stats.note.empty() || !d.timer_tag().empty()) {
        std::shared_lock<std::shared_mutex> g(g_capture_lock);
        step_eager(t);
        return;
    }
    if (m_step_exec) {
        if (step_replay(t)) return;
        drop_step_graph();  // rolled back: the step runs eagerly and is captured again
    }
    StepCapture &cap = d.capture();
    cap = StepCapture{};
    cap.mode = StepCapture::kRecord;
    try {
        std::shared_lock<std::shared_mutex> g(g_capture_lock);
        step_eager(t);
    } catch (...) {
        cap.mode = StepCapture::kOff;
        throw;
    }
    cap.mode = StepCapture::kOff;
    if (cap.failed || cap.counts.empty() || cap.counts.size() > static_cast<size_t>(StepCapture::kMaxLoops))
        return;  // capture after a step whose loops all ended by their tests
    std::unique_lock<std::shared_mutex> g(g_capture_lock);
    step_capture();
}

void Solver::set_step_graph(bool on, int margin) {
    drop_step_graph();
    m_graph_on = on;
    m_graph_margin = margin;
    m_graph_stats = GraphStats{};
    if (!on || m_loop_log) return;
    Domain &d = *m_domain;
    ARTSS_HIP(hipMalloc(&m_loop_log, sizeof(LoopLog) * StepCapture::kMaxLoops));
    ARTSS_HIP(hipHostMalloc(&m_loop_log_host, sizeof(LoopLog) * StepCapture::kMaxLoops));
    const size_t n = static_cast<size_t>(d.level(0).size);
    auto fields = m_fields->all();
    std::vector<double *> state;
    for (auto &kv : fields) state.push_back(kv.second->data);
    // the warm start's previous solution (kept in place while a step is captured)
    if (!m_p_prev) {
        const LevelDesc &L = d.level(0);
        ARTSS_HIP(hipMalloc(&m_p_prev, sizeof(double) * n));
        ARTSS_HIP(hipMalloc(&m_p_start, sizeof(double) * n));
        ARTSS_HIP(hipMalloc(&m_warm_partial, sizeof(double) * static_cast<size_t>(kk::blocks_inner(L))));
        ARTSS_HIP(hipMalloc(&m_warm_sum, sizeof(double)));
    }
    state.push_back(m_p_prev);
    for (size_t q = 0; q < state.size(); q++) {
        double *c = nullptr;
        ARTSS_HIP(hipMalloc(&c, sizeof(double) * n));
        m_rollback.push_back(c);
    }
    m_state_count = static_cast<int>(state.size());
    ARTSS_HIP(hipMalloc(&m_state_dev, sizeof(double *) * state.size()));
    ARTSS_HIP(hipMalloc(&m_rollback_dev, sizeof(double *) * state.size()));
    ARTSS_HIP(hipMemcpy(m_state_dev, state.data(), sizeof(double *) * state.size(), hipMemcpyHostToDevice));
    ARTSS_HIP(hipMemcpy(m_rollback_dev, m_rollback.data(), sizeof(double *) * state.size(), hipMemcpyHostToDevice));
}

void Solver::drop_step_graph() {
    if (m_step_exec) {
        (void)hipStreamSynchronize(m_domain->stream());
        (void)hipGraphExecDestroy(m_step_exec);
        m_step_exec = nullptr;
    }
    m_graph_stats.active = false;
}

// records the step launched by the next replays: the fields copied aside,
// then the step with each residual-target loop on a budget
void Solver::step_capture() {
    Domain &d = *m_domain;
    StepCapture &cap = d.capture();
    std::vector<int> budgets = cap.counts;
    for (int &b : budgets) b = std::max(0, b + m_graph_margin);
    m_loop_kind = cap.kinds;
    const size_t n = static_cast<size_t>(d.level(0).size);
    cap.mode = StepCapture::kCapture;
    cap.next = 0;
    cap.budgets = budgets;
    cap.log = m_loop_log;
    hipGraph_t graph = nullptr;
    // the pass that records the graph runs the step's host code: the
    // counters of the step just taken stay
    const StepCounters taken = m_counters;
    ARTSS_HIP(hipStreamBeginCapture(d.stream(), hipStreamCaptureModeRelaxed));
    try {
        k::copy_many(d.stream(), m_rollback_dev, m_state_dev, m_state_count, static_cast<long long>(n));
        step_eager(0.0);  // time enters the captured sources through their device factors
        m_counters = taken;
    } catch (const std::exception &e) {
        m_counters = taken;
        (void)hipStreamEndCapture(d.stream(), &graph);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        cap.mode = StepCapture::kOff;
        m_graph_stats.note = std::string("not captured: ") + e.what();
        return;
    }
    cap.mode = StepCapture::kOff;
    ARTSS_HIP(hipStreamEndCapture(d.stream(), &graph));
    if (cap.next != static_cast<int>(budgets.size())) {
        (void)hipGraphDestroy(graph);
        m_graph_stats.note = "not captured: the captured step has other loops than the recorded one";
        return;
    }
    const hipError_t e = hipGraphInstantiate(&m_step_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    ARTSS_HIP(e);
    m_graph_stats.captures++;
    m_graph_stats.active = true;
}

// one replay; false (with the fields rolled back) when a loop would have
// needed more than its budget
bool Solver::step_replay(double t) {
    Domain &d = *m_domain;
    hipStream_t s = d.stream();
    if (m_gauss) m_gauss->set_time(t);
    if (m_gauss_con) m_gauss_con->set_time(t);
    const size_t loops = m_loop_kind.size();
    ARTSS_HIP(hipGraphLaunch(m_step_exec, s));
    ARTSS_HIP(hipMemcpyAsync(m_loop_log_host, m_loop_log, sizeof(LoopLog) * loops, hipMemcpyDeviceToHost, s));
    ARTSS_HIP(hipStreamSynchronize(s));
    bool ok = true;
    for (size_t q = 0; q < loops; q++) ok = ok && m_loop_log_host[q].ok != 0;
    if (!ok) {
        const size_t n = static_cast<size_t>(d.level(0).size);
        k::copy_many(s, m_state_dev, m_rollback_dev, m_state_count, static_cast<long long>(n));
        ARTSS_HIP(hipStreamSynchronize(s));
        m_graph_stats.rollbacks++;
        return false;
    }
    m_counters = StepCounters{};
    int dif = 0;
    for (size_t q = 0; q < loops; q++) {
        const LoopLog &l = m_loop_log_host[q];
        if (m_loop_kind[q] == StepCapture::kPressure) {
            m_counters.vcycles = l.count;
            m_counters.residual = l.value;
            m_counters.initial_residual = l.start;
            m_counters.converged = true;
        } else if (dif < 4) {
            m_counters.diffusion[dif++] = l.count;
        }
    }
    if (m_warm.mode != 0) m_warm_history++;
    m_graph_stats.replays++;
    return true;
}

void Solver::step_eager(double t) {
    m_counters = StepCounters{};
    switch (m_kind) {
        case 0: do_step_ns(t); break;
        case 1: do_step_nstempturb(t); break;
        case 2: pressure_step(t); break;
        case 3: do_step_nstemp(t); break;
        case 4: do_step_nsturb(t); break;
        case 9:
            do_step_nstemp(t);
            do_step_concentration();
            break;
        case 10:
            do_step_nstempturb(t);
            do_step_concentration();
            break;
        default: do_step_transport(); break;
    }
    update_sources(t);
    m_fields->update_data();
}

int Solver::snapshot(double t) {
    Domain &d = *m_domain;
    const size_t n = static_cast<size_t>(d.level(0).size);
    auto snap = std::make_unique<Snapshot>();
    snap->t = t;
    auto fields = m_fields->all();
    try {
        for (auto &kv : fields) {
            double *q = nullptr;
            ARTSS_HIP(hipMalloc(&q, sizeof(double) * n));
            snap->data.push_back(q);
            ARTSS_HIP(hipMemcpyAsync(q, kv.second->data, sizeof(double) * n, hipMemcpyDeviceToDevice, d.stream()));
        }
        if (m_p_prev) {
            double *q = nullptr;
            ARTSS_HIP(hipMalloc(&q, sizeof(double) * n));
            snap->data.push_back(q);
            ARTSS_HIP(hipMemcpyAsync(q, m_p_prev, sizeof(double) * n, hipMemcpyDeviceToDevice, d.stream()));
        }
    } catch (...) {
        (void)hipStreamSynchronize(d.stream());
        for (double *q : snap->data) (void)hipFree(q);
        throw;
    }
    snap->warm_history = m_warm_history;
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
    for (size_t i = 0; i < m_snapshots.size(); i++)
        if (!m_snapshots[i]) {
            m_snapshots[i] = std::move(snap);
            return static_cast<int>(i);
        }
    m_snapshots.push_back(std::move(snap));
    return static_cast<int>(m_snapshots.size()) - 1;
}

double Solver::restore(int slot) {
    if (slot < 0 || slot >= static_cast<int>(m_snapshots.size()) || !m_snapshots[static_cast<size_t>(slot)])
        throw config_error("no snapshot in slot " + std::to_string(slot));
    Domain &d = *m_domain;
    const size_t n = static_cast<size_t>(d.level(0).size);
    const Snapshot &snap = *m_snapshots[static_cast<size_t>(slot)];
    auto fields = m_fields->all();
    for (size_t i = 0; i < fields.size(); i++)
        ARTSS_HIP(hipMemcpyAsync(fields[i].second->data, snap.data[i], sizeof(double) * n, hipMemcpyDeviceToDevice,
                                 d.stream()));
    if (m_p_prev && snap.data.size() > fields.size()) {
        ARTSS_HIP(hipMemcpyAsync(m_p_prev, snap.data[fields.size()], sizeof(double) * n, hipMemcpyDeviceToDevice,
                                 d.stream()));
        m_warm_history = snap.warm_history;
    } else {
        m_warm_history = 0;  // no previous solution in the snapshot: the next solve starts plainly
    }
    m_warm_sample = false;
    // the projection start keeps no increments across a restore: its next
    // start uses the one increment the snapshot holds; a captured step is
    // recorded again
    proj_reset();
    drop_step_graph();
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
    return snap.t;
}

void Solver::drop_snapshot(int slot) {
    if (slot < 0 || slot >= static_cast<int>(m_snapshots.size()) || !m_snapshots[static_cast<size_t>(slot)]) return;
    (void)hipStreamSynchronize(m_domain->stream());
    for (double *q : m_snapshots[static_cast<size_t>(slot)]->data) (void)hipFree(q);
    m_snapshots[static_cast<size_t>(slot)].reset();
}

namespace {
constexpr char kCheckpointMagic[8] = {'A', 'R', 'T', 'S', 'S', 'H', 'I', 'P'};
constexpr uint32_t kCheckpointVersion = 1;
}  // namespace

void Solver::save(const std::string &path, double t) {
    Domain &d = *m_domain;
    const LevelDesc &L = d.level(0);
    const size_t n = static_cast<size_t>(L.size);
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw config_error("cannot write checkpoint '" + path + "'");
    auto fields = m_fields->all();
    const bool warm = m_p_prev != nullptr;
    std::vector<std::string> names;
    for (auto &kv : fields) names.push_back(kv.first);
    if (warm) names.push_back("p_prev");
    const uint32_t header[6] = {kCheckpointVersion, static_cast<uint32_t>(L.N[0]), static_cast<uint32_t>(L.N[1]),
                                static_cast<uint32_t>(L.N[2]), static_cast<uint32_t>(names.size()),
                                static_cast<uint32_t>(m_warm_history)};
    const double times[2] = {t, d.settings().dt};
    out.write(kCheckpointMagic, sizeof(kCheckpointMagic));
    out.write(reinterpret_cast<const char *>(header), sizeof(header));
    out.write(reinterpret_cast<const char *>(times), sizeof(times));
    for (const std::string &nm : names) {
        const uint32_t len = static_cast<uint32_t>(nm.size());
        out.write(reinterpret_cast<const char *>(&len), sizeof(len));
        out.write(nm.data(), len);
    }
    std::vector<double> host(n);
    // the copies run on the solver's stream, behind any step still in flight
    for (size_t i = 0; i < names.size(); i++) {
        const double *src = i < fields.size() ? fields[i].second->data : m_p_prev;
        ARTSS_HIP(hipMemcpyAsync(host.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, d.stream()));
        ARTSS_HIP(hipStreamSynchronize(d.stream()));
        out.write(reinterpret_cast<const char *>(host.data()), static_cast<std::streamsize>(sizeof(double) * n));
    }
    if (!out) throw config_error("writing checkpoint '" + path + "' failed");
}

double Solver::load(const std::string &path) {
    Domain &d = *m_domain;
    const LevelDesc &L = d.level(0);
    const size_t n = static_cast<size_t>(L.size);
    std::ifstream in(path, std::ios::binary);
    if (!in) throw config_error("cannot read checkpoint '" + path + "'");
    char magic[8];
    uint32_t header[6];
    double times[2];
    in.read(magic, sizeof(magic));
    in.read(reinterpret_cast<char *>(header), sizeof(header));
    in.read(reinterpret_cast<char *>(times), sizeof(times));
    if (!in || std::memcmp(magic, kCheckpointMagic, sizeof(magic)) != 0 || header[0] != kCheckpointVersion)
        throw config_error("'" + path + "' is not a checkpoint of this format");
    if (header[1] != static_cast<uint32_t>(L.N[0]) || header[2] != static_cast<uint32_t>(L.N[1]) ||
        header[3] != static_cast<uint32_t>(L.N[2]))
        throw config_error("checkpoint '" + path + "' is for another grid");
    std::vector<std::string> names(header[4]);
    for (std::string &nm : names) {
        uint32_t len = 0;
        in.read(reinterpret_cast<char *>(&len), sizeof(len));
        if (!in || len > 64) throw config_error("corrupt checkpoint '" + path + "'");
        nm.resize(len);
        in.read(&nm[0], len);
    }
    std::vector<double> host(n);
    bool have_prev = false;
    for (const std::string &nm : names) {
        in.read(reinterpret_cast<char *>(host.data()), static_cast<std::streamsize>(sizeof(double) * n));
        if (!in) throw config_error("checkpoint '" + path + "' is truncated");
        double *dst = nullptr;
        if (nm == "p_prev") {
            if (!m_p_prev) continue;  // no warm start set up here: nothing to continue
            dst = m_p_prev;
            have_prev = true;
        } else {
            Field *f = m_fields->by_name(nm);
            if (f == nullptr) throw config_error("checkpoint '" + path + "' holds an unknown field '" + nm + "'");
            dst = f->data;
        }
        // ordered after the kernels still in flight on the solver's stream;
        // the synchronisation keeps `host` alive until the copy is done
        ARTSS_HIP(hipMemcpyAsync(dst, host.data(), sizeof(double) * n, hipMemcpyHostToDevice, d.stream()));
        ARTSS_HIP(hipStreamSynchronize(d.stream()));
    }
    m_warm_history = have_prev ? static_cast<int>(header[5]) : 0;
    proj_reset();
    drop_step_graph();
    m_warm_sample = false;
    return times[0];
}

int Solver::write_dat(const std::string &dir) {
    Domain &d = *m_domain;
    static const char *names[NUMBER_OF_FIELD_TYPES] = {"rho", "u", "v", "w", "p", "T", "nu"};
    const std::vector<uint64_t> obstacle_cells = d.obstacle_list(0);
    int files = 0;
    std::vector<double> host(static_cast<size_t>(d.level(0).size));
    for (int ft = 0; ft < NUMBER_OF_FIELD_TYPES; ft++) {
        if (!d.domain_bc(0, ft).active) continue;  // BoundaryDataController::get_used_fields
        Field *f = m_fields->by_name(names[ft]);
        if (f == nullptr) continue;  // no such field on this path (rho, nu)
        ARTSS_HIP(hipMemcpyAsync(host.data(), f->data, sizeof(double) * host.size(), hipMemcpyDeviceToHost,
                                 d.stream()));
        ARTSS_HIP(hipStreamSynchronize(d.stream()));
        {
            std::ofstream out(dir + "/" + names[ft] + ".dat");
            if (!out) throw std::invalid_argument("cannot write " + dir + "/" + names[ft] + ".dat");
            for (double x : host) out << x << std::endl;
        }
        files++;
        if (!obstacle_cells.empty()) {
            std::ofstream out(dir + "/" + names[ft] + "_obstacle.dat");
            if (!out) throw std::invalid_argument("cannot write " + dir + "/" + names[ft] + "_obstacle.dat");
            for (uint64_t c : obstacle_cells) out << c << ";" << host[c] << std::endl;
            files++;
        }
    }
    return files;
}

}  // namespace artss
