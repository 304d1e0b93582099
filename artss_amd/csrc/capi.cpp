// C ABI (include/artss_hip.h) over the C++ host side.
#include "../../include/artss_hip.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "ensemble.h"
#include "residual_static.h"
#include "solver.h"
#include "unet.h"

using namespace artss;

struct artss_domain {
    Domain *dom = nullptr;
    std::unique_ptr<Domain> owned;
    std::unique_ptr<GaussFunction> gauss;
};

struct artss_mg {
    artss_domain *d;
    std::unique_ptr<VCycleMG> mg;
};

struct artss_unet {
    artss_domain *d;
    std::unique_ptr<UNet> net;
};

struct artss_solver {
    std::unique_ptr<Solver> solver;
    artss_domain view;
};

struct artss_ensemble {
    std::unique_ptr<Ensemble> ensemble;
};

namespace {

thread_local std::string g_error;

template <typename F>
int guard(F &&f) {
    try {
        g_error.clear();
        return f();
    } catch (const config_error &e) {
        g_error = e.what();
        return ARTSS_ERR_CONFIG;
    } catch (const hip_error &e) {
        g_error = e.what();
        return ARTSS_ERR_HIP;
    } catch (const std::invalid_argument &e) {
        g_error = e.what();
        return ARTSS_ERR_ARGUMENT;
    } catch (const std::exception &e) {
        g_error = e.what();
        return ARTSS_ERR_INTERNAL;
    }
}

void require(bool ok, const char *what) {
    if (!ok) throw std::invalid_argument(what);
}

Domain &dom(const artss_domain *d) {
    require(d && d->dom, "null domain");
    return *d->dom;
}

void check_level(const Domain &D, int level) { require(level >= 0 && level <= D.levels(), "level out of range"); }
void check_type(int ft) { require(ft >= 0 && ft < NUMBER_OF_FIELD_TYPES, "field type out of range"); }

void finish(Domain &D) {
    ARTSS_HIP(hipGetLastError());
    ARTSS_HIP(hipStreamSynchronize(D.stream()));
}

// status of a completed step: a residual-target pressure solve that missed
// its target is reported, with the residual, after the step is done
int step_status(const Solver &s, long step) {
    const StepCounters &c = s.counters();
    if (c.converged) return ARTSS_OK;
    g_error = "step " + std::to_string(step) + ": pressure residual " + std::to_string(c.residual) +
              " above the target after " + std::to_string(c.vcycles) + " iterations";
    return ARTSS_ERR_NOT_CONVERGED;
}

std::vector<uint64_t> patch_list(const Domain &D, int level, int patch) {
    // Domain::boundary_cells, Domain.cpp:94-172
    const LevelDesc &L = D.level(level);
    const int axis = patch / 2;
    const int ax1 = axis == 0 ? 1 : 0, ax2 = axis == 2 ? 1 : 2;
    std::vector<uint64_t> out;
    size_t c[3];
    c[axis] = (patch % 2) ? L.N[axis] - 1 : 0;
    for (int k = 0; k < L.N[ax2]; k++)
        for (int j = 0; j < L.N[ax1]; j++) {
            c[ax2] = k;
            c[ax1] = j;
            out.push_back(c[0] + static_cast<size_t>(L.N[0]) * (c[1] + static_cast<size_t>(L.N[1]) * c[2]));
        }
    return out;
}

}  // namespace

extern "C" {

const char *artss_last_error(void) { return g_error.c_str(); }
int artss_version(void) { return 1; }

int artss_domain_create(const char *xml_text, int device, artss_domain **out) {
    return guard([&] {
        require(xml_text && out, "null argument");
        auto d = std::make_unique<artss_domain>();
        d->owned = std::make_unique<Domain>(parse_settings(xml_text), device);
        d->dom = d->owned.get();
        *out = d.release();
        return ARTSS_OK;
    });
}

int artss_domain_destroy(artss_domain *d) {
    return guard([&] {
        require(d && d->owned, "not an owned domain");
        delete d;
        return ARTSS_OK;
    });
}

int artss_domain_levels(const artss_domain *d, int *levels) {
    return guard([&] {
        require(levels, "null argument");
        *levels = dom(d).levels();
        return ARTSS_OK;
    });
}

int artss_domain_level_info(const artss_domain *d, int level, artss_level_info *info) {
    return guard([&] {
        const Domain &D = dom(d);
        check_level(D, level);
        require(info, "null argument");
        const LevelDesc &L = D.level(level);
        for (int a = 0; a < 3; a++) {
            info->N[a] = L.N[a];
            info->n[a] = L.n[a];
            info->h[a] = L.h[a];
        }
        info->size = L.size;
        info->inner_cells = static_cast<long long>(D.inner_cells(level));
        info->n_obstacles = L.n_obst;
        return ARTSS_OK;
    });
}

int artss_domain_list(const artss_domain *d, int which, int level, int obstacle, uint64_t *out, size_t capacity,
                      size_t *count) {
    return guard([&] {
        const Domain &D = dom(d);
        check_level(D, level);
        require(count, "null count");
        std::vector<uint64_t> v;
        if (which == 0) v = D.inner_list(level);
        else if (which == 1) v = D.domain_list(level);
        else if (which == 2) v = D.obstacle_list(level);
        else if (which >= 3 && which <= 8) v = patch_list(D, level, which - 3);
        else if (which >= 10 && which <= 15) {
            require(obstacle >= 0 && obstacle < D.obstacle_count(), "obstacle out of range");
            v = D.obstacle(level, obstacle).face[which - 10];
        } else {
            require(false, "unknown list");
        }
        *count = v.size();
        if (out) {
            require(capacity >= v.size(), "list capacity too small");
            if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(uint64_t));
        }
        return ARTSS_OK;
    });
}

int artss_domain_synchronize(artss_domain *d) {
    return guard([&] {
        finish(dom(d));
        return ARTSS_OK;
    });
}

int artss_field_alloc(artss_domain *d, int level, double **dev) {
    return guard([&] {
        Domain &D = dom(d);
        check_level(D, level);
        require(dev, "null argument");
        *dev = D.alloc_field(level);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_field_free(artss_domain *d, double *dev) {
    return guard([&] {
        dom(d).free_field(dev);
        return ARTSS_OK;
    });
}

int artss_field_upload(artss_domain *d, double *dev, const double *host, size_t n) {
    return guard([&] {
        Domain &D = dom(d);
        require(dev && host, "null argument");
        ARTSS_HIP(hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, D.stream()));
        finish(D);
        return ARTSS_OK;
    });
}

int artss_field_download(artss_domain *d, double *host, const double *dev, size_t n) {
    return guard([&] {
        Domain &D = dom(d);
        require(dev && host, "null argument");
        ARTSS_HIP(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, D.stream()));
        finish(D);
        return ARTSS_OK;
    });
}

int artss_apply_boundary(artss_domain *d, double *field, int field_type, int level) {
    return guard([&] {
        Domain &D = dom(d);
        check_level(D, level);
        check_type(field_type);
        require(field, "null field");
        D.apply_boundary(field, field_type, level);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_sl_advect(artss_domain *d, double *out, const double *in, const double *u, const double *v,
                    const double *w, int out_type) {
    return guard([&] {
        Domain &D = dom(d);
        check_type(out_type);
        require(out && in && u && v && w, "null field");
        Field fo(D, out_type, 0, out), fi(D, out_type, 0, const_cast<double *>(in));
        Field fu(D, U, 0, const_cast<double *>(u)), fv(D, V, 0, const_cast<double *>(v)),
            fw(D, W, 0, const_cast<double *>(w));
        SLAdvect(D).advect(fo, fi, fu, fv, fw, false);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_diffuse(artss_domain *d, const char *method, double *out, double *in, const double *b, double Dc,
                  const double *ev, int field_type, int max_iter, double tol_res, double w, int *iterations,
                  double *residual) {
    return guard([&] {
        Domain &D = dom(d);
        check_type(field_type);
        require(method && out && in && b, "null argument");
        DiffusionSettings s;
        s.type = method;
        s.max_iter = max_iter;
        s.tol_res = tol_res;
        s.w = w;
        std::unique_ptr<IDiffusion> dif = make_diffusion(D, s);
        Field fo(D, field_type, 0, out), fi(D, field_type, 0, in), fb(D, field_type, 0, const_cast<double *>(b));
        if (ev) {
            Field fe(D, NU, 0, const_cast<double *>(ev));
            dif->diffuse(fo, fi, fb, Dc, fe, false);
        } else {
            dif->diffuse(fo, fi, fb, Dc, false);
        }
        finish(D);
        if (iterations) *iterations = dif->last_iterations();
        if (residual) *residual = dif->last_residual();
        return ARTSS_OK;
    });
}

int artss_jacobi_step(artss_domain *d, int level, double *out, const double *in, const double *b, double ax,
                      double ay, double az, double beta, double dsign, double w) {
    return guard([&] {
        Domain &D = dom(d);
        check_level(D, level);
        require(out && in && b, "null field");
        k::jacobi_step_mg(D.stream(), D.level(level), out, in, b, ax, ay, az, beta, dsign, w);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_cgs_step(artss_domain *d, int level, double *out, const double *b, double ax, double ay, double az,
                   double beta, double dsign, double w) {
    return guard([&] {
        Domain &D = dom(d);
        check_level(D, level);
        require(out && b, "null field");
        for (int color = 0; color < 2; color++)
            k::cgs_color(D.stream(), D.level(level), color, out, b, ax, ay, az, beta, dsign, w, 0.0, nullptr, 0.0);
        finish(D);
        return ARTSS_OK;
    });
}

namespace {
struct PressureOps : IPressure {
    explicit PressureOps(Domain &d) : IPressure(d) {}
    void pressure(Field &, const Field &, double, bool) override {}
};
}  // namespace

int artss_divergence(artss_domain *d, double *out, const double *x, const double *y, const double *z) {
    return guard([&] {
        Domain &D = dom(d);
        require(out && x && y && z, "null field");
        Field fo(D, P, 0, out), fx(D, U, 0, const_cast<double *>(x)), fy(D, V, 0, const_cast<double *>(y)),
            fz(D, W, 0, const_cast<double *>(z));
        PressureOps(D).divergence(fo, fx, fy, fz, false);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_projection(artss_domain *d, double *ou, double *ov, double *ow, const double *iu, const double *iv,
                     const double *iw, const double *ip) {
    return guard([&] {
        Domain &D = dom(d);
        require(ou && ov && ow && iu && iv && iw && ip, "null field");
        Field a(D, U, 0, ou), b(D, V, 0, ov), c(D, W, 0, ow);
        Field e(D, U, 0, const_cast<double *>(iu)), f(D, V, 0, const_cast<double *>(iv)),
            g(D, W, 0, const_cast<double *>(iw)), h(D, P, 0, const_cast<double *>(ip));
        PressureOps(D).projection(a, b, c, e, f, g, h, false);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_smagorinsky(artss_domain *d, double *ev, const double *u, const double *v, const double *w, double cs,
                      int ev_type) {
    return guard([&] {
        Domain &D = dom(d);
        check_type(ev_type);
        require(ev && u && v && w, "null field");
        Field fe(D, ev_type, 0, ev), fu(D, U, 0, const_cast<double *>(u)), fv(D, V, 0, const_cast<double *>(v)),
            fw(D, W, 0, const_cast<double *>(w));
        ConstSmagorinsky(D, cs).calc_turbulent_viscosity(fe, fu, fv, fw, false);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_add_source(artss_domain *d, double *out, const double *s, int field_type) {
    return guard([&] {
        Domain &D = dom(d);
        check_type(field_type);
        require(out && s, "null field");
        Field fo(D, field_type, 0, out), fs(D, field_type, 0, const_cast<double *>(s));
        ExplicitEulerSource(D, {}).add_source(fo, fs, false);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_buoyancy_force(artss_domain *d, double *out, const double *T, const double *Ta) {
    return guard([&] {
        Domain &D = dom(d);
        require(out && T && Ta, "null field");
        Field fo(D, V, 0, out), ft(D, artss::T, 0, const_cast<double *>(T)), fa(D, artss::T, 0, const_cast<double *>(Ta));
        ISource::buoyancy_force(fo, ft, fa, D.settings().beta, D.settings().g, false);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_gauss_source(artss_domain *d, double *out, double t) {
    return guard([&] {
        Domain &D = dom(d);
        require(out, "null field");
        require(D.settings().temp_fct == "Gauss", "the XML input has no Gauss temperature source");
        if (!d->gauss) d->gauss = std::make_unique<GaussFunction>(D, D.settings().gauss);
        Field fo(D, artss::T, 0, out);
        d->gauss->update_source(fo, t);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_mg_create(artss_domain *d, artss_mg **out) {
    return guard([&] {
        Domain &D = dom(d);
        require(out, "null argument");
        require(D.settings().has_pressure, "the XML input has no pressure solver");
        auto m = std::make_unique<artss_mg>();
        m->d = d;
        m->mg = std::make_unique<VCycleMG>(D, D.settings().pressure);
        finish(D);
        *out = m.release();
        return ARTSS_OK;
    });
}

int artss_mg_destroy(artss_mg *mg) {
    return guard([&] {
        require(mg, "null argument");
        delete mg;
        return ARTSS_OK;
    });
}

int artss_mg_pressure(artss_mg *mg, double *p, const double *rhs, double t, int *cycles) {
    return guard([&] {
        require(mg && p && rhs, "null argument");
        Domain &D = dom(mg->d);
        Field fp(D, P, 0, p), fb(D, P, 0, const_cast<double *>(rhs));
        mg->mg->pressure(fp, fb, t, false);
        finish(D);
        if (cycles) *cycles = mg->mg->last_cycles();
        return ARTSS_OK;
    });
}

int artss_mg_vcycle(artss_mg *mg, double *p, const double *rhs) {
    return guard([&] {
        require(mg && p && rhs, "null argument");
        Domain &D = dom(mg->d);
        Field fp(D, P, 0, p), fb(D, P, 0, const_cast<double *>(rhs));
        mg->mg->vcycle(fp, fb);
        finish(D);
        return ARTSS_OK;
    });
}

int artss_mg_set_schedule(artss_mg *mg, int mode, long long tail_cells) {
    return guard([&] {
        require(mg && tail_cells >= 0 && mode >= 0 && mode <= 2, "bad argument");
        mg->mg->set_schedule(mode != 0, tail_cells);
        mg->mg->set_graphs(mode == 2);
        finish(dom(mg->d));
        return ARTSS_OK;
    });
}

int artss_mg_tail_info(artss_mg *mg, int *first, int *lds_bytes) {
    return guard([&] {
        require(mg && first, "null argument");
        *first = mg->mg->tail_first();
        if (lds_bytes) *lds_bytes = mg->mg->tail_lds_bytes();
        return ARTSS_OK;
    });
}

int artss_mg_residual(artss_mg *mg, const double *p, const double *rhs, double *residual) {
    return guard([&] {
        require(mg && p && rhs && residual, "null argument");
        Domain &D = dom(mg->d);
        Field fp(D, P, 0, const_cast<double *>(p)), fb(D, P, 0, const_cast<double *>(rhs));
        *residual = mg->mg->residual(fp, fb);
        return ARTSS_OK;
    });
}

int artss_mg_residual_static(artss_mg *mg, const double *p, const double *rhs, double *residual) {
    return guard([&] {
        require(mg && p && rhs && residual, "null argument");
        *residual = residual_static(dom(mg->d), p, rhs);
        return ARTSS_OK;
    });
}

int artss_solver_create(const char *xml_text, int device, artss_solver **out) {
    return guard([&] {
        require(xml_text && out, "null argument");
        auto s = std::make_unique<artss_solver>();
        s->solver = std::make_unique<Solver>(std::make_unique<Domain>(parse_settings(xml_text), device));
        s->view.dom = &s->solver->domain();
        *out = s.release();
        return ARTSS_OK;
    });
}

int artss_solver_destroy(artss_solver *s) {
    return guard([&] {
        require(s, "null argument");
        delete s;
        return ARTSS_OK;
    });
}

int artss_solver_domain(artss_solver *s, artss_domain **d) {
    return guard([&] {
        require(s && d, "null argument");
        *d = &s->view;
        return ARTSS_OK;
    });
}

int artss_solver_step(artss_solver *s, double t) {
    return guard([&] {
        require(s, "null argument");
        s->solver->step(t);
        finish(s->solver->domain());
        return step_status(*s->solver, std::lround(t / s->solver->domain().settings().dt));
    });
}

int artss_solver_run(artss_solver *s, long first_step, long count, double *seconds) {
    return guard([&] {
        require(s && count >= 0 && first_step >= 1, "bad argument");
        Domain &D = s->solver->domain();
        const double dt = D.settings().dt;
        finish(D);
        const auto t0 = std::chrono::steady_clock::now();
        long missed = 0;
        for (long n = first_step; n < first_step + count; n++) {
            s->solver->step(static_cast<double>(n) * dt);
            if (!s->solver->counters().converged && missed == 0) missed = n;
        }
        finish(D);
        const auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        if (missed) {
            g_error = "step " + std::to_string(missed) + " and possibly later ones: pressure residual above the target";
            return ARTSS_ERR_NOT_CONVERGED;
        }
        return ARTSS_OK;
    });
}

int artss_solver_field(artss_solver *s, const char *name, double **dev, size_t *n) {
    return guard([&] {
        require(s && name && dev, "null argument");
        Field *f = s->solver->fields().by_name(name);
        require(f != nullptr, "unknown field name");
        *dev = f->data;
        if (n) *n = static_cast<size_t>(f->get_size());
        return ARTSS_OK;
    });
}

int artss_solver_counters(artss_solver *s, int *diffusion4, int *vcycles, double *residual) {
    return guard([&] {
        require(s, "null argument");
        const StepCounters &c = s->solver->counters();
        if (diffusion4)
            for (int i = 0; i < 4; i++) diffusion4[i] = c.diffusion[i];
        if (vcycles) *vcycles = c.vcycles;
        if (residual) *residual = c.residual;
        return ARTSS_OK;
    });
}

int artss_solver_set_pressure_mode(artss_solver *s, int to_tolerance, double tolerance, int max_cycles) {
    return guard([&] {
        require(s && max_cycles >= 0, "bad argument");
        PressureMode m = s->solver->pressure_mode();
        m.to_tolerance = to_tolerance != 0;
        m.tolerance = tolerance;
        m.max_cycles = max_cycles;
        s->solver->set_pressure_mode(m);
        return ARTSS_OK;
    });
}

int artss_solver_write_dat(artss_solver *s, const char *directory, int *files) {
    return guard([&] {
        require(s && directory, "null argument");
        const int n = s->solver->write_dat(directory);
        if (files) *files = n;
        return ARTSS_OK;
    });
}

int artss_solver_set_pressure_method(artss_solver *s, int method, int coarse_max_iter) {
    return guard([&] {
        require(s && method >= 0 && method <= 2 && coarse_max_iter >= 0, "bad argument");
        PressureMode m = s->solver->pressure_mode();
        m.method = method;
        m.coarse_max_iter = coarse_max_iter;
        s->solver->set_pressure_mode(m);
        return ARTSS_OK;
    });
}

int artss_solver_snapshot(artss_solver *s, double t, int *slot) {
    return guard([&] {
        require(s && slot, "null argument");
        *slot = s->solver->snapshot(t);
        return ARTSS_OK;
    });
}

int artss_solver_restore(artss_solver *s, int slot, double *t) {
    return guard([&] {
        require(s, "null argument");
        const double tt = s->solver->restore(slot);
        if (t) *t = tt;
        return ARTSS_OK;
    });
}

int artss_solver_drop_snapshot(artss_solver *s, int slot) {
    return guard([&] {
        require(s, "null argument");
        s->solver->drop_snapshot(slot);
        return ARTSS_OK;
    });
}

int artss_solver_save(artss_solver *s, const char *path, double t) {
    return guard([&] {
        require(s && path, "null argument");
        s->solver->save(path, t);
        return ARTSS_OK;
    });
}

int artss_solver_load(artss_solver *s, const char *path, double *t) {
    return guard([&] {
        require(s && path, "null argument");
        const double tt = s->solver->load(path);
        if (t) *t = tt;
        return ARTSS_OK;
    });
}

int artss_ensemble_create(int threads, artss_ensemble **out) {
    return guard([&] {
        require(out && threads >= 1, "bad argument");
        auto e = std::make_unique<artss_ensemble>();
        e->ensemble = std::make_unique<Ensemble>(threads);
        *out = e.release();
        return ARTSS_OK;
    });
}

int artss_ensemble_destroy(artss_ensemble *e) {
    return guard([&] {
        require(e, "null argument");
        delete e;
        return ARTSS_OK;
    });
}

int artss_ensemble_add(artss_ensemble *e, artss_solver *member) {
    return guard([&] {
        require(e && member, "null argument");
        e->ensemble->add(member->solver.get());
        return ARTSS_OK;
    });
}

int artss_ensemble_step(artss_ensemble *e, long step) {
    return guard([&] {
        require(e, "null argument");
        e->ensemble->step(step);
        for (size_t m = 0; m < e->ensemble->size(); m++)
            if (!e->ensemble->member(m).counters().converged) {
                g_error = "member " + std::to_string(m) + ", step " + std::to_string(step) +
                          ": pressure residual above the target";
                return ARTSS_ERR_NOT_CONVERGED;
            }
        return ARTSS_OK;
    });
}

int artss_solver_set_warm_start(artss_solver *s, int mode, artss_unet *net, int record) {
    return guard([&] {
        require(s, "null argument");
        if (net) require(dom(net->d).stream() == s->solver->domain().stream(),
                         "the network must be created on the solver's domain");
        WarmStart w;
        w.mode = mode;
        w.net = net ? net->net.get() : nullptr;
        w.record = record != 0;
        s->solver->set_warm_start(w);
        return ARTSS_OK;
    });
}

int artss_solver_set_step_graph(artss_solver *s, int on, int margin) {
    return guard([&] {
        require(s && margin >= -1000 && margin <= 1000, "bad argument");
        s->solver->set_step_graph(on != 0, margin);
        return ARTSS_OK;
    });
}

int artss_solver_step_graph_stats(artss_solver *s, long long *replays, long long *rollbacks, long long *captures,
                                  int *active, const char **note) {
    return guard([&] {
        require(s && replays && rollbacks && captures && active && note, "null argument");
        const Solver::GraphStats &g = s->solver->graph_stats();
        *replays = g.replays;
        *rollbacks = g.rollbacks;
        *captures = g.captures;
        *active = g.active ? 1 : 0;
        *note = g.note.c_str();
        return ARTSS_OK;
    });
}

int artss_solver_set_warm_start_depth(artss_solver *s, int mode, int depth) {
    return guard([&] {
        require(s, "null argument");
        WarmStart w;
        w.mode = mode;
        w.depth = depth;
        s->solver->set_warm_start(w);
        return ARTSS_OK;
    });
}

int artss_solver_warm_sample(artss_solver *s, float *x, float *t, int *available) {
    return guard([&] {
        require(s && x && t && available, "null argument");
        *available = s->solver->warm_sample(x, t) ? 1 : 0;
        return ARTSS_OK;
    });
}

int artss_solver_initial_residual(artss_solver *s, double *residual) {
    return guard([&] {
        require(s && residual, "null argument");
        *residual = s->solver->counters().initial_residual;
        return ARTSS_OK;
    });
}

int artss_tuned_cells_per_thread(long long level_size) { return artss::k::tuned_kz(level_size); }

int artss_solver_set_mixed_tail(artss_solver *s, int n_relax) {
    return guard([&] {
        require(s && n_relax >= 0, "bad argument");
        PressureMode m = s->solver->pressure_mode();
        m.mixed_tail = n_relax;
        s->solver->set_pressure_mode(m);
        return ARTSS_OK;
    });
}

int artss_solver_fallbacks(artss_solver *s, long long *steps) {
    return guard([&] {
        require(s && steps, "null argument");
        *steps = s->solver->fallbacks();
        return ARTSS_OK;
    });
}

int artss_solver_set_coarse_tolerance(artss_solver *s, double coarse_tol) {
    return guard([&] {
        require(s && std::isfinite(coarse_tol) && coarse_tol < 1.0, "bad argument");
        PressureMode m = s->solver->pressure_mode();
        m.coarse_tol = coarse_tol;
        s->solver->set_pressure_mode(m);
        return ARTSS_OK;
    });
}

int artss_solver_set_coarse_solve(artss_solver *s, int direct) {
    return guard([&] {
        require(s && (direct == 0 || direct == 1), "bad argument");
        PressureMode m = s->solver->pressure_mode();
        m.coarse_direct = direct == 1;
        s->solver->set_pressure_mode(m);
        return ARTSS_OK;
    });
}

int artss_solver_set_mixed_sweeps(artss_solver *s, int pre, int post, double weight) {
    return guard([&] {
        require(s && pre >= 0 && post >= 0 && weight >= 0.0 && weight < 2.0, "bad argument");
        PressureMode m = s->solver->pressure_mode();
        m.mixed_pre = pre;
        m.mixed_post = post;
        m.mixed_weight = weight;
        s->solver->set_pressure_mode(m);
        return ARTSS_OK;
    });
}

int artss_solver_set_mg_schedule(artss_solver *s, int mode, long long tail_cells) {
    return guard([&] {
        require(s && tail_cells >= 0 && mode >= 0 && mode <= 2, "bad argument");
        s->solver->invalidate_step_graph();
        s->solver->pressure_solver().set_schedule(mode != 0, tail_cells);
        s->solver->pressure_solver().set_graphs(mode == 2);
        finish(s->solver->domain());
        return ARTSS_OK;
    });
}

int artss_buffer_alloc(artss_domain *d, size_t bytes, void **dev) {
    return guard([&] {
        require(dev && bytes > 0, "bad argument");
        dom(d);
        void *p = nullptr;
        ARTSS_HIP(hipMalloc(&p, bytes));
        ARTSS_HIP(hipMemset(p, 0, bytes));
        *dev = p;
        return ARTSS_OK;
    });
}

int artss_buffer_free(artss_domain *d, void *dev) {
    return guard([&] {
        dom(d);
        if (dev) ARTSS_HIP(hipFree(dev));
        return ARTSS_OK;
    });
}

int artss_buffer_upload(artss_domain *d, void *dev, const void *host, size_t bytes) {
    return guard([&] {
        require(dev && host, "null argument");
        Domain &D = dom(d);
        ARTSS_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, D.stream()));
        ARTSS_HIP(hipStreamSynchronize(D.stream()));
        return ARTSS_OK;
    });
}

int artss_buffer_download(artss_domain *d, void *host, const void *dev, size_t bytes) {
    return guard([&] {
        require(dev && host, "null argument");
        Domain &D = dom(d);
        ARTSS_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, D.stream()));
        ARTSS_HIP(hipStreamSynchronize(D.stream()));
        return ARTSS_OK;
    });
}

int artss_unet_create(artss_domain *d, unsigned long long seed, artss_unet **out) {
    return guard([&] {
        require(out, "null argument");
        Domain &D = dom(d);
        auto u = std::make_unique<artss_unet>();
        u->d = d;
        u->net = std::make_unique<UNet>(D.stream(), seed);
        *out = u.release();
        return ARTSS_OK;
    });
}

int artss_unet_destroy(artss_unet *u) {
    return guard([&] {
        require(u, "null argument");
        delete u;
        return ARTSS_OK;
    });
}

int artss_unet_param_count(size_t *n) {
    return guard([&] {
        require(n, "null argument");
        *n = UNet::param_count();
        return ARTSS_OK;
    });
}

int artss_unet_set_params(artss_unet *u, const float *host) {
    return guard([&] {
        require(u && host, "null argument");
        u->net->set_params(host);
        return ARTSS_OK;
    });
}

int artss_unet_get_params(artss_unet *u, float *host) {
    return guard([&] {
        require(u && host, "null argument");
        u->net->get_params(host);
        return ARTSS_OK;
    });
}

int artss_unet_forward(artss_unet *u, int X, int Y, int Z, const float *x, float *y) {
    return guard([&] {
        require(u && x && y, "null argument");
        u->net->forward(X, Y, Z, x, y);
        return ARTSS_OK;
    });
}

int artss_unet_gradients(artss_unet *u, int X, int Y, int Z, const float *x, const float *target, float *grads,
                         float *loss) {
    return guard([&] {
        require(u && x && target, "null argument");
        const float l = u->net->gradients(X, Y, Z, x, target, grads);
        if (loss) *loss = l;
        return ARTSS_OK;
    });
}

int artss_unet_adam_step(artss_unet *u, float lr) {
    return guard([&] {
        require(u && lr > 0, "bad argument");
        u->net->adam_step(lr);
        return ARTSS_OK;
    });
}

int artss_profile_arm(artss_domain *d, const char *tag) {
    return guard([&] {
        require(tag, "null tag");
        dom(d).timer_arm(tag);
        return ARTSS_OK;
    });
}

int artss_profile_read(artss_domain *d, long long *launches, double *total_ms) {
    return guard([&] {
        require(launches && total_ms, "null argument");
        dom(d).timer_read(launches, total_ms);
        return ARTSS_OK;
    });
}

}  // extern "C"
