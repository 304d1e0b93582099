#include "operators.h"

#include "dyn_smagorinsky.h"

#include <cmath>

namespace artss {

namespace {
void maybe_sync(Domain &d, bool sync) {
    if (sync) ARTSS_HIP(hipStreamSynchronize(d.stream()));
}
}  // namespace

// IPressure.cpp:20-110
void IPressure::divergence(Field &out, const Field &in_x, const Field &in_y, const Field &in_z, bool sync) {
    k::divergence(m_domain.stream(), m_domain.level(0), out.data, in_x.data, in_y.data, in_z.data);
    maybe_sync(m_domain, sync);
}

// IPressure.cpp:125-172
void IPressure::projection(Field &out_u, Field &out_v, Field &out_w, const Field &in_u, const Field &in_v,
                           const Field &in_w, const Field &in_p, bool sync) {
    k::projection(m_domain.stream(), m_domain.level(0), out_u.data, out_v.data, out_w.data, in_u.data, in_v.data,
                  in_w.data, in_p.data);
    m_domain.apply_boundary(out_u.data, out_u.get_type(), 0);
    m_domain.apply_boundary(out_v.data, out_v.get_type(), 0);
    m_domain.apply_boundary(out_w.data, out_w.get_type(), 0);
    maybe_sync(m_domain, sync);
}

// ISource.cpp:26-52
void ISource::buoyancy_force(Field &out, const Field &in, const Field &in_a, double beta, double g, bool sync) {
    Domain &d = out.domain();
    k::buoyancy(d.stream(), d.level(0), out.data, in.data, in_a.data, beta, g);
    maybe_sync(d, sync);
}

void ISource::dissipate(Field &out, const Field &in_u, const Field &in_v, const Field &in_w, bool sync) {
    Domain &d = out.domain();
    k::dissipate(d.stream(), d.level(0), out.data, in_u.data, in_v.data, in_w.data, d.settings().nu,
                 d.settings().dt);
    d.apply_boundary(out.data, out.get_type(), 0);
    maybe_sync(d, sync);
}

void SLAdvect::advect(Field &out, const Field &in, const Field &u, const Field &v, const Field &w, bool sync) {
    m_domain.timed("sl_advect", 0, [&] {
        k::sl_advect(m_domain.stream(), m_domain.level(0), m_domain.settings().dt, out.data, in.data, u.data,
                     v.data, w.data);
    });
    m_domain.apply_boundary(out.data, out.get_type(), 0);
    maybe_sync(m_domain, sync);
}

// JacobiDiffuse.cpp:31-110 (EV == nullptr) and :123-203.
// Iterate while the update size exceeds tol_res; `in` follows `out` on the
// domain list after every sweep, its obstacle cells keep their entry values.
void JacobiDiffuse::run(Field &out, const Field &in, const Field &b, double D, const Field *EV) {
    Domain &dm = m_domain;
    const LevelDesc &L = dm.level(0);
    const double dt = dm.settings().dt;
    const double rdx = 1. / L.h[0], rdy = 1. / L.h[1], rdz = 1. / L.h[2];
    const double ax = D * dt * rdx * rdx, ay = D * dt * rdy * rdy, az = D * dt * rdz * rdz;
    const double rb = (1. + 2. * (ax + ay + az));
    const int n_part = k::partial_count(L);
    StepCapture &cap = dm.capture();
    if (cap.mode == StepCapture::kCapture) throw capture_unsupported("the Jacobi diffusion loop");
    if (cap.mode == StepCapture::kRecord) cap.failed = true;
    int it = 0;
    double res = 1.;
    while (res > m_settings.tol_res && it < m_settings.max_iter) {
        dm.timed("jacobi_diffuse", 0, [&] {
            k::jacobi_diffuse_step(dm.stream(), L, out.data, in.data, b.data, ax, ay, az, rb, D,
                                   EV ? EV->data : nullptr, dt, m_settings.w, dm.partial_buffer());
        });
        dm.apply_boundary(out.data, out.get_type(), 0);
        res = std::sqrt(dm.reduce(dm.partial_buffer(), n_part));
        it++;
        k::copy_domain(dm.stream(), L, in.data, out.data);
    }
    if (it % 2 != 0) k::copy_domain(dm.stream(), L, in.data, out.data);
    m_iterations = it;
    m_residual = res;
}

void JacobiDiffuse::diffuse(Field &out, const Field &in, const Field &b, double D, bool sync) {
    run(out, in, b, D, nullptr);
    maybe_sync(m_domain, sync);
}
void JacobiDiffuse::diffuse(Field &out, const Field &in, const Field &b, double D, const Field &EV, bool sync) {
    run(out, in, b, D, &EV);
    maybe_sync(m_domain, sync);
}

// ExplicitDiffuse.cpp:27-47 (no iteration: one step, then the fill)
void ExplicitDiffuse::run(Field &out, const Field &in, double D, const Field *EV) {
    Domain &dm = m_domain;
    dm.timed("explicit_diffuse", 0, [&] {
        k::explicit_diffuse(dm.stream(), dm.level(0), out.data, in.data, D, EV ? EV->data : nullptr,
                            dm.settings().dt);
    });
    dm.apply_boundary(out.data, out.get_type(), 0);
    m_iterations = 1;
    m_residual = 0.0;
}
void ExplicitDiffuse::diffuse(Field &out, const Field &in, const Field &, double D, bool sync) {
    run(out, in, D, nullptr);
    maybe_sync(m_domain, sync);
}
void ExplicitDiffuse::diffuse(Field &out, const Field &in, const Field &, double D, const Field &EV, bool sync) {
    run(out, in, D, &EV);
    maybe_sync(m_domain, sync);
}

// ColoredGaussSeidelDiffuse.cpp:38-173: red then black in place on `out`,
// `in` holds the previous iterate (whole-array copy) for the update norm.
void ColoredGaussSeidelDiffuse::run(Field &out, const Field &in, const Field &b, double D, const Field *EV) {
    Domain &dm = m_domain;
    const LevelDesc &L = dm.level(0);
    const double dt = dm.settings().dt;
    const double rdx = 1. / L.h[0], rdy = 1. / L.h[1], rdz = 1. / L.h[2];
    const double ax = D * dt * rdx * rdx, ay = D * dt * rdy * rdy, az = D * dt * rdz * rdz;
    const double rb = (1. + 2. * (ax + ay + az));
    const double beta = 1. / rb;
    const double *ev = EV ? EV->data : nullptr;
    // The loop test runs on the device (k::IterGate): as many iterations as
    // the previous call took are enqueued at once, then one at a time; the
    // kernels of iterations past the test's stop do nothing, so the fields
    // are those of the reference's loop, which tests after every iteration.
    hipStream_t s = dm.stream();
    if (!m_gate) {
        ARTSS_HIP(hipMalloc(&m_gate, sizeof(k::IterGate)));
        ARTSS_HIP(hipHostMalloc(&m_gate_host, sizeof(k::IterGate)));
    }
    const int *stop = &m_gate->done;
    const int ft = out.get_type();
    if (dm.bc_idempotent(0, ft)) {
        const BCView V = dm.bc_view(0, ft);
        if (k::cgs_iter_partials(L) <= dm.partial_capacity()) {
            const k::CgsCoef c{ax, ay, az, beta, rb, 1.0, m_settings.w, D, dt, rdx, rdy, rdz};
            run_one_pass(out, in, b, ev, V, c);
            return;
        }
    }
    StepCapture &cap = dm.capture();
    if (cap.mode == StepCapture::kCapture)
        throw capture_unsupported("the colored Gauss-Seidel loop with a fill that is not idempotent");
    if (cap.mode == StepCapture::kRecord) cap.failed = true;
    const int n_part = k::partial_count(L);
    out.copy_data(in);
    k::gate_open(s, m_gate, m_settings.tol_res, m_settings.max_iter);
    int launched = 0;
    while (true) {
        const int batch = launched < m_batch_hint ? m_batch_hint - launched : 1;
        for (int q = 0; q < batch; q++) {
            k::copy(s, in.data, out.data, in.get_size(), stop);
            for (int color = 0; color < 2; color++)
                dm.timed("cgs_diffuse", 0, [&] {
                    k::cgs_color(s, L, color, out.data, b.data, ax, ay, az, beta, 1.0, m_settings.w, D, ev, dt,
                                 stop);
                });
            dm.apply_boundary(out.data, out.get_type(), 0, stop);
            k::update_norm_partials(s, L, out.data, in.data, rb, D, ev, dt, dm.partial_buffer(), stop);
            k::reduce_gate(s, dm.partial_buffer(), n_part, m_gate);
        }
        launched += batch;
        ARTSS_HIP(hipMemcpyAsync(m_gate_host, m_gate, sizeof(k::IterGate), hipMemcpyDeviceToHost, s));
        ARTSS_HIP(hipStreamSynchronize(s));
        if (m_gate_host->done) break;
    }
    m_iterations = m_gate_host->count;
    m_residual = m_iterations > 0 ? std::sqrt(m_gate_host->sum) : 1.;
    m_batch_hint = m_iterations;
}

// The same loop with each iteration as one out-of-place kernel
// (k::cgs_iter: both colours, the fill and the norm partials).  The
// reference copies in <- out before every iteration and updates out in
// place, so after iteration q `out` holds x_q and `in` x_(q-1).  Here
// iteration 1 reads `in` (the reference's out equals it then) and writes
// `out`, iteration 2 reads `out` and writes `in`, and so on: after an odd
// count both arrays hold what the reference leaves in them, after an even
// count their contents are exchanged once at the end.
void ColoredGaussSeidelDiffuse::run_one_pass(Field &out, const Field &in, const Field &b, const double *ev,
                                             const BCView &V, const k::CgsCoef &c) {
    Domain &dm = m_domain;
    const LevelDesc &L = dm.level(0);
    hipStream_t s = dm.stream();
    const int *stop = &m_gate->done;
    const int n_part = k::cgs_iter_partials(L);
    k::gate_open(s, m_gate, m_settings.tol_res, m_settings.max_iter);
    StepCapture &cap = dm.capture();
    if (cap.mode == StepCapture::kCapture) {
        // a fixed budget behind the gate; the arrays' final exchange and the
        // outcome are settled on the device
        const int slot = cap.next++;
        if (slot >= StepCapture::kMaxLoops || slot >= static_cast<int>(cap.budgets.size()))
            throw capture_unsupported("more residual-target loops than the recorded step had");
        for (int q = 0; q < cap.budgets[static_cast<size_t>(slot)]; q++) {
            const bool odd = (q & 1) == 0;
            k::cgs_iter(s, L, V, c, ev, b.data, odd ? in.data : out.data, odd ? out.data : in.data,
                        dm.partial_buffer(), stop);
            k::reduce_gate(s, dm.partial_buffer(), n_part, m_gate);
        }
        k::alternate_finish(s, m_gate, out.data, in.data, out.get_size());
        k::log_gate(s, m_gate, cap.log + slot);
        return;
    }
    int launched = 0;
    while (true) {
        const int batch = launched < m_batch_hint ? m_batch_hint - launched : 1;
        for (int q = 0; q < batch; q++) {
            const bool odd = ((launched + q) & 1) == 0;  // iteration number launched + q + 1
            const double *src = odd ? in.data : out.data;
            double *dst = odd ? out.data : in.data;
            dm.timed("cgs_diffuse", 0, [&] {
                k::cgs_iter(s, L, V, c, ev, b.data, src, dst, dm.partial_buffer(), stop);
            });
            k::reduce_gate(s, dm.partial_buffer(), n_part, m_gate);
        }
        launched += batch;
        ARTSS_HIP(hipMemcpyAsync(m_gate_host, m_gate, sizeof(k::IterGate), hipMemcpyDeviceToHost, s));
        ARTSS_HIP(hipStreamSynchronize(s));
        if (m_gate_host->done) break;
    }
    m_iterations = m_gate_host->count;
    m_residual = m_iterations > 0 ? std::sqrt(m_gate_host->sum) : 1.;
    m_batch_hint = m_iterations;
    if (m_iterations == 0) out.copy_data(in);
    else if (m_iterations % 2 == 0) k::swap_arrays(s, out.data, in.data, out.get_size());
    if (cap.mode == StepCapture::kRecord) {
        cap.counts.push_back(m_iterations);
        cap.kinds.push_back(StepCapture::kDiffusion);
    }
}

ColoredGaussSeidelDiffuse::~ColoredGaussSeidelDiffuse() {
    if (m_gate) {
        (void)hipStreamSynchronize(m_domain.stream());
        (void)hipFree(m_gate);
        (void)hipHostFree(m_gate_host);
    }
}

void ColoredGaussSeidelDiffuse::diffuse(Field &out, const Field &in, const Field &b, double D, bool sync) {
    run(out, in, b, D, nullptr);
    maybe_sync(m_domain, sync);
}
void ColoredGaussSeidelDiffuse::diffuse(Field &out, const Field &in, const Field &b, double D, const Field &EV,
                                        bool sync) {
    run(out, in, b, D, &EV);
    maybe_sync(m_domain, sync);
}

// ConstSmagorinsky.cpp:30-87; Cs*Cs*delta*delta is folded on the host in the
// reference's left-to-right order before the per-cell multiply by S_bar.
void ConstSmagorinsky::calc_turbulent_viscosity(Field &ev, const Field &u, const Field &v, const Field &w,
                                                bool sync) {
    const LevelDesc &L = m_domain.level(0);
    const double delta_s = std::cbrt(L.h[0] * L.h[1] * L.h[2]);
    const double cdd = m_cs * m_cs * delta_s * delta_s;
    k::smagorinsky(m_domain.stream(), L, ev.data, u.data, v.data, w.data, cdd);
    m_domain.apply_boundary(ev.data, ev.get_type(), 0);
    maybe_sync(m_domain, sync);
}

void ExplicitEulerSource::add_source(Field &out_x, Field &out_y, Field &out_z, const Field &S_x, const Field &S_y,
                                     const Field &S_z, bool sync) {
    const double dt = m_domain.settings().dt;
    Field *o[3] = {&out_x, &out_y, &out_z};
    const Field *s[3] = {&S_x, &S_y, &S_z};
    for (int a = 0; a < 3; a++) {
        bool on = false;
        for (int d : m_dir) on = on || d == a;
        if (!on) continue;
        k::add_source(m_domain.stream(), m_domain.level(0), o[a]->data, s[a]->data, dt);
        m_domain.apply_boundary(o[a]->data, o[a]->get_type(), 0);
    }
    maybe_sync(m_domain, sync);
}

void ExplicitEulerSource::add_source(Field &out, const Field &S, bool sync) {
    k::add_source(m_domain.stream(), m_domain.level(0), out.data, S.data, m_domain.settings().dt);
    m_domain.apply_boundary(out.data, out.get_type(), 0);
    maybe_sync(m_domain, sync);
}

GaussFunction::GaussFunction(Domain &d, const GaussSettings &g)
    : m_domain(d), m_settings(g), m_spatial(d, RHO) {
    const std::vector<double> host = d.gauss_spatial(g);
    ARTSS_HIP(hipMemcpyAsync(m_spatial.data, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                             d.stream()));
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
    ARTSS_HIP(hipHostMalloc(&m_tv_host, sizeof(double)));
    ARTSS_HIP(hipMalloc(&m_tv_dev, sizeof(double)));
    *m_tv_host = 0.0;
}

GaussFunction::~GaussFunction() {
    (void)hipStreamSynchronize(m_domain.stream());
    (void)hipHostFree(m_tv_host);
    (void)hipFree(m_tv_dev);
}

void GaussFunction::set_time(double t_cur) { *m_tv_host = std::tanh(t_cur / m_settings.tau); }

// GaussFunction.cpp:27-40, 121-123: out = spatial * tanh(t / tau)
void GaussFunction::update_source(Field &out, double t_cur) {
    hipStream_t s = m_domain.stream();
    if (m_domain.capture().mode == StepCapture::kCapture) {
        ARTSS_HIP(hipMemcpyAsync(m_tv_dev, m_tv_host, sizeof(double), hipMemcpyHostToDevice, s));
        k::scale_to_dev(s, out.data, m_spatial.data, m_tv_dev, out.get_size());
        return;
    }
    const double tv = std::tanh(t_cur / m_settings.tau);
    k::scale_to(s, out.data, m_spatial.data, tv, out.get_size());
}

BuoyancyMMS::BuoyancyMMS(Domain &d) : m_domain(d), m_source(d, RHO) {
    const Settings &s = d.settings();
    const LevelDesc &L = d.level(0);
    std::vector<double> host(static_cast<size_t>(L.size), 0.0);
    const double X1 = s.X[0], Y1 = s.Y[0];
    const double dx = L.h[0], dy = L.h[1];
    const double rbeta = 1. / s.beta;
    const double rg = 1. / s.g;
    const double c_nu = 2 * s.nu * M_PI * M_PI - 1;
    const double c_kappa = 2 * s.kappa * M_PI * M_PI - 1;
    const size_t Nx = static_cast<size_t>(L.N[0]), Ny = static_cast<size_t>(L.N[1]);
    for (uint64_t idx : d.domain_list(0)) {
        const size_t k = idx / (Nx * Ny);
        const size_t j = (idx - k * Nx * Ny) / Nx;
        const size_t i = idx - k * Nx * Ny - j * Nx;
        // xi(i, X1, dx) = X1 + (i - 0.5) * dx (utility/GlobalMacrosTypes.h:76)
        const double x = X1 + (static_cast<double>(i) - 0.5) * dx;
        const double y = Y1 + (static_cast<double>(j) - 0.5) * dy;
        host[idx] = s.rhoa * rbeta * rg * 2 * c_nu * c_kappa * std::sin(M_PI * (x + y));
    }
    ARTSS_HIP(hipMemcpyAsync(m_source.data, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                             d.stream()));
    ARTSS_HIP(hipStreamSynchronize(d.stream()));
}

// BuoyancyMMS.cpp:58-61: out = source * exp(-t) over the whole array
void BuoyancyMMS::update_source(Field &out, double t_cur) {
    if (m_domain.capture().mode == StepCapture::kCapture)
        throw capture_unsupported("the manufactured buoyancy source (its factor exp(-t) is a kernel argument)");
    k::scale_to(m_domain.stream(), out.data, m_source.data, std::exp(-t_cur), out.get_size());
}

std::unique_ptr<IAdvection> make_advection(Domain &d, const std::string &type) {
    if (type == "SemiLagrangian") return std::make_unique<SLAdvect>(d);
    throw config_error("advection method '" + type + "' is not implemented");
}

std::unique_ptr<IDiffusion> make_diffusion(Domain &d, const DiffusionSettings &s) {
    if (s.type == "Jacobi") return std::make_unique<JacobiDiffuse>(d, s);
    if (s.type == "ColoredGaussSeidel") return std::make_unique<ColoredGaussSeidelDiffuse>(d, s);
    if (s.type == "Explicit") return std::make_unique<ExplicitDiffuse>(d);
    throw config_error("diffusion method '" + s.type + "' is not implemented");
}

std::unique_ptr<ITurbulence> make_turbulence(Domain &d, const Settings &s) {
    if (s.turbulence_type == "ConstSmagorinsky") return std::make_unique<ConstSmagorinsky>(d, s.cs);
    if (s.turbulence_type == "DynamicSmagorinsky") return std::make_unique<DynamicSmagorinsky>(d);
    throw config_error("turbulence method '" + s.turbulence_type + "' is not implemented");
}

}  // namespace artss
