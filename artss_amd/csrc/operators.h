// Operator interfaces of the reference (src/interfaces/*.h) and their
// HIP-backed implementations.  Method names, argument order and meaning are
// the reference's so a solver written against IAdvection / IDiffusion /
// IPressure / ITurbulence / ISource drives these unchanged; the domain is
// passed at construction instead of being fetched from a process-wide
// singleton, so several domains (ensemble members) can live in one process.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "field.h"
#include "settings.h"

namespace artss {

// interfaces/IAdvection.h:16
class IAdvection {
  public:
    virtual ~IAdvection() = default;
    virtual void advect(Field &out, const Field &in, const Field &u_vel, const Field &v_vel, const Field &w_vel,
                        bool sync) = 0;
};

// interfaces/IDiffusion.h:16-19
class IDiffusion {
  public:
    virtual ~IDiffusion() = default;
    virtual void diffuse(Field &out, const Field &in, const Field &b, double D, bool sync) = 0;
    virtual void diffuse(Field &out, const Field &in, const Field &b, double D, const Field &EV, bool sync) = 0;
    int last_iterations() const { return m_iterations; }
    double last_residual() const { return m_residual; }

  protected:
    int m_iterations = 0;
    double m_residual = 0.0;
};

// interfaces/IPressure.h:17-24 (divergence/projection are IPressure.cpp)
class IPressure {
  public:
    explicit IPressure(Domain &d) : m_domain(d) {}
    virtual ~IPressure() = default;
    virtual void pressure(Field &out, const Field &b, double t, bool sync) = 0;
    void divergence(Field &out, const Field &in_x, const Field &in_y, const Field &in_z, bool sync);
    void projection(Field &out_u, Field &out_v, Field &out_w, const Field &in_u, const Field &in_v,
                    const Field &in_w, const Field &in_p, bool sync);

  protected:
    Domain &m_domain;
};

// interfaces/ITurbulence.h:18-22
class ITurbulence {
  public:
    virtual ~ITurbulence() = default;
    virtual void calc_turbulent_viscosity(Field &ev, const Field &in_u, const Field &in_v, const Field &in_w,
                                          bool sync) = 0;
    virtual void explicit_filtering(Field &out, const Field &in, bool sync) = 0;
};

// interfaces/ISource.h:18-29
class ISource {
  public:
    virtual ~ISource() = default;
    virtual void add_source(Field &out_x, Field &out_y, Field &out_z, const Field &S_x, const Field &S_y,
                            const Field &S_z, bool sync) = 0;
    virtual void add_source(Field &out, const Field &S, bool sync) = 0;
    static void buoyancy_force(Field &out, const Field &in, const Field &in_a, double beta, double g, bool sync);
    // ISource.cpp:63-121: out += dt * nu * (viscous dissipation of u, v, w), then the fill
    static void dissipate(Field &out, const Field &in_u, const Field &in_v, const Field &in_w, bool sync);
};

// ---------------------------------------------------------------- implementations

// advection/SLAdvect.cpp:34-170
class SLAdvect : public IAdvection {
  public:
    explicit SLAdvect(Domain &d) : m_domain(d) {}
    void advect(Field &out, const Field &in, const Field &u_vel, const Field &v_vel, const Field &w_vel,
                bool sync) override;

  private:
    Domain &m_domain;
};

// diffusion/JacobiDiffuse.cpp:31-203
class JacobiDiffuse : public IDiffusion {
  public:
    JacobiDiffuse(Domain &d, const DiffusionSettings &s) : m_domain(d), m_settings(s) {}
    void diffuse(Field &out, const Field &in, const Field &b, double D, bool sync) override;
    void diffuse(Field &out, const Field &in, const Field &b, double D, const Field &EV, bool sync) override;

  private:
    void run(Field &out, const Field &in, const Field &b, double D, const Field *EV);
    Domain &m_domain;
    DiffusionSettings m_settings;
};

// diffusion/ExplicitDiffuse.cpp:27-47: one explicit Euler step and the fill
class ExplicitDiffuse : public IDiffusion {
  public:
    explicit ExplicitDiffuse(Domain &d) : m_domain(d) {}
    void diffuse(Field &out, const Field &in, const Field &b, double D, bool sync) override;
    void diffuse(Field &out, const Field &in, const Field &b, double D, const Field &EV, bool sync) override;

  private:
    void run(Field &out, const Field &in, double D, const Field *EV);
    Domain &m_domain;
};

// diffusion/ColoredGaussSeidelDiffuse.cpp:38-173
class ColoredGaussSeidelDiffuse : public IDiffusion {
  public:
    ColoredGaussSeidelDiffuse(Domain &d, const DiffusionSettings &s) : m_domain(d), m_settings(s) {}
    ~ColoredGaussSeidelDiffuse() override;
    ColoredGaussSeidelDiffuse(const ColoredGaussSeidelDiffuse &) = delete;
    ColoredGaussSeidelDiffuse &operator=(const ColoredGaussSeidelDiffuse &) = delete;
    void diffuse(Field &out, const Field &in, const Field &b, double D, bool sync) override;
    void diffuse(Field &out, const Field &in, const Field &b, double D, const Field &EV, bool sync) override;

  private:
    void run(Field &out, const Field &in, const Field &b, double D, const Field *EV);
    void run_one_pass(Field &out, const Field &in, const Field &b, const double *ev, const BCView &V,
                      const k::CgsCoef &c);
    Domain &m_domain;
    DiffusionSettings m_settings;
    k::IterGate *m_gate = nullptr;       // device loop state
    k::IterGate *m_gate_host = nullptr;  // pinned
    int m_batch_hint = 0;                // iterations of the previous call, enqueued at once
};

// turbulence/ConstSmagorinsky.cpp:30-97
class ConstSmagorinsky : public ITurbulence {
  public:
    ConstSmagorinsky(Domain &d, double cs) : m_domain(d), m_cs(cs) {}
    void calc_turbulent_viscosity(Field &ev, const Field &in_u, const Field &in_v, const Field &in_w,
                                  bool sync) override;
    void explicit_filtering(Field &, const Field &, bool) override {}

  private:
    Domain &m_domain;
    double m_cs;
};

// source/ExplicitEulerSource.cpp:25-104
class ExplicitEulerSource : public ISource {
  public:
    ExplicitEulerSource(Domain &d, std::vector<int> dir) : m_domain(d), m_dir(std::move(dir)) {}
    void add_source(Field &out_x, Field &out_y, Field &out_z, const Field &S_x, const Field &S_y, const Field &S_z,
                    bool sync) override;
    void add_source(Field &out, const Field &S, bool sync) override;

  private:
    Domain &m_domain;
    std::vector<int> m_dir;
};

// source/GaussFunction.cpp:16-123: spatial part built once on the host,
// time ramp tanh(t/tau) applied on the device each update.
class GaussFunction {
  public:
    GaussFunction(Domain &d, const GaussSettings &g);
    ~GaussFunction();
    GaussFunction(const GaussFunction &) = delete;
    GaussFunction &operator=(const GaussFunction &) = delete;
    void update_source(Field &out, double t_cur);
    const Field &spatial() const { return m_spatial; }
    // a captured step reads the time factor tanh(t / tau) from device memory,
    // copied there from pinned memory that set_time writes before each replay
    void set_time(double t_cur);

  private:
    Domain &m_domain;
    GaussSettings m_settings;
    Field m_spatial;
    double *m_tv_host = nullptr;  // pinned
    double *m_tv_dev = nullptr;
};

// source/BuoyancyMMS.cpp:11-69 (temp_fct="Buoyancy", SolverSelection.cpp:148-149):
// the manufactured energy source rhoa/(beta g) 2 c_nu c_kappa sin(pi (x + y))
// on the level-0 domain list, built once on the host with libm in the
// reference's expression order; update_source scales it by exp(-t).
class BuoyancyMMS {
  public:
    explicit BuoyancyMMS(Domain &d);
    // not capturable (its factor exp(-t) is a kernel argument)
    void update_source(Field &out, double t_cur);

  private:
    Domain &m_domain;
    Field m_source;
};

// SolverSelection (solver/SolverSelection.cpp:33-213): same type strings
std::unique_ptr<IAdvection> make_advection(Domain &d, const std::string &type);
std::unique_ptr<IDiffusion> make_diffusion(Domain &d, const DiffusionSettings &s);
std::unique_ptr<ITurbulence> make_turbulence(Domain &d, const Settings &s);

}  // namespace artss
