#include "settings.h"

#include <algorithm>
#include <cctype>

#include <cerrno>
#include <cstdlib>
#include <memory>

#include "xml.h"

namespace artss {
namespace {

using Map = std::map<std::string, std::string>;

// Settings.cpp map_parameter_section / map_parameter_line: attributes plus
// the text of each child element.
Map values_of(const xml::Element *e) {
    Map m;
    if (!e) return m;
    for (const auto &a : e->attributes) m[a.first] = a.second;
    for (const auto &c : e->children)
        if (c->has_text) m[c->name] = c->text;
    return m;
}

const xml::Element *section(const xml::Element *head, const std::string &name, const std::string &ctx) {
    const xml::Element *e = head ? head->child(name) : nullptr;
    if (!e) throw config_error("section " + ctx + "/" + name + " is required.");
    return e;
}

double to_real(const std::string &s, const std::string &key, const std::string &ctx) {
    errno = 0;
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || errno == ERANGE) throw config_error("Value " + key + " is for " + ctx + " required.");
    return v;
}
long to_int(const std::string &s, const std::string &key, const std::string &ctx) {
    errno = 0;
    char *end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || errno == ERANGE) throw config_error("Value " + key + " is for " + ctx + " required.");
    return v;
}
double req_real(const Map &m, const std::string &k, const std::string &ctx) {
    auto it = m.find(k);
    if (it == m.end()) throw config_error("Value " + k + " is for " + ctx + " required.");
    return to_real(it->second, k, ctx);
}
double opt_real(const Map &m, const std::string &k, double dflt) {
    auto it = m.find(k);
    return it == m.end() ? dflt : to_real(it->second, k, "");
}
int req_int(const Map &m, const std::string &k, const std::string &ctx) {
    auto it = m.find(k);
    if (it == m.end()) throw config_error("Value " + k + " is for " + ctx + " required.");
    return static_cast<int>(to_int(it->second, k, ctx));
}
std::string req_str(const Map &m, const std::string &k, const std::string &ctx) {
    auto it = m.find(k);
    if (it == m.end()) throw config_error("Value " + k + " is for " + ctx + " required.");
    return it->second;
}
bool req_bool(const Map &m, const std::string &k, const std::string &ctx) { return req_str(m, k, ctx) == "Yes"; }
bool opt_bool(const Map &m, const std::string &k, bool dflt) {
    auto it = m.find(k);
    return it == m.end() ? dflt : it->second == "Yes";
}

// Settings.cpp:166-186; Functions.cpp:740-748 rejects a relative range
// outside [0, 1]
RandomSettings random_settings(const xml::Element *e, const std::string &ctx) {
    const Map v = values_of(e);
    RandomSettings r;
    r.absolute = req_bool(v, "absolute", ctx);
    r.custom_seed = req_bool(v, "custom_seed", ctx);
    auto seed = v.find("seed");
    if (r.custom_seed && seed == v.end()) throw config_error("Value seed is for " + ctx + " required.");
    if (seed != v.end()) {
        errno = 0;
        char *end = nullptr;
        r.seed = std::strtoull(seed->second.c_str(), &end, 10);
        if (end == seed->second.c_str() || errno == ERANGE) throw config_error("Value seed is for " + ctx + " required.");
    }
    r.step_size = req_bool(v, "custom_steps", ctx) ? req_real(v, "step_size", ctx) : opt_real(v, "step_size", 1.0);
    r.range = req_real(v, "range", ctx);
    if (!r.absolute && (r.range > 1 || r.range < 0))
        throw config_error("range for relative noise has to be between [0:1]");
    return r;
}

std::vector<std::string> split(const std::string &s, char d) {
    std::vector<std::string> out;
    size_t p = 0;
    for (;;) {
        size_t e = s.find(d, p);
        out.push_back(s.substr(p, e == std::string::npos ? std::string::npos : e - p));
        if (e == std::string::npos) break;
        p = e + 1;
    }
    return out;
}

int match(const std::vector<std::string> &names, const std::string &s, const std::string &what) {
    for (size_t i = 0; i < names.size(); i++)
        if (names[i] == s) return static_cast<int>(i);
    throw config_error("unknown " + what + " '" + s + "'");
}
const std::vector<std::string> kFieldNames = {"rho", "u", "v", "w", "p", "T", "nu"};
const std::vector<std::string> kPatchNames = {"left", "right", "bottom", "top", "front", "back"};
const std::vector<std::string> kBCNames = {"neumann", "dirichlet", "periodic"};
const std::vector<std::string> kAxisNames = {"x", "y", "z"};

// Settings.cpp:716-718 with Mapping::match_axis (Mapping.cpp:17-23): each
// comma-separated entry is matched case-insensitively; an entry that names
// no axis ("xyz" in several of the reference's inputs) becomes UNKNOWN_AXIS,
// which no caller ever looks for (SolverController.cpp:239-281,
// ExplicitEulerSource.cpp:40-70 test for X, Y and Z only), so it is dropped.
// Only membership matters to the callers, so each axis is kept once.
std::vector<int> axes(const std::string &s) {
    std::vector<int> out;
    for (const auto &a : split(s, ',')) {
        std::string up = a;
        for (char &c : up) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        for (size_t i = 0; i < kAxisNames.size(); i++)
            if (kAxisNames[i] == up && std::find(out.begin(), out.end(), static_cast<int>(i)) == out.end())
                out.push_back(static_cast<int>(i));
    }
    return out;
}

// Settings.cpp:372-389
BoundarySpec parse_boundary(const xml::Element *e, const std::string &ctx) {
    Map v = values_of(e);
    BoundarySpec b;
    b.bc = match(kBCNames, req_str(v, "type", ctx), "boundary condition");
    if (b.bc != PERIODIC) b.value = req_real(v, "value", ctx);
    for (const auto &f : split(req_str(v, "field", ctx), ',')) b.fields.push_back(match(kFieldNames, f, "field"));
    for (const auto &p : split(req_str(v, "patch", ctx), ',')) b.patches.push_back(match(kPatchNames, p, "patch"));
    return b;
}

// Settings.cpp:752-769
GaussSettings parse_gauss(const Map &v, const std::string &c) {
    GaussSettings g;
    g.hrr = req_real(v, "HRR", c);
    g.cp = req_real(v, "cp", c);
    const char *ax[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; a++) {
        g.pos[a] = req_real(v, std::string(ax[a]) + "0", c);
        g.sigma[a] = req_real(v, std::string("sigma_") + ax[a], c);
    }
    g.tau = req_real(v, "tau", c);
    return g;
}

// Settings.cpp:771-781
CubeSettings parse_cube(const Map &v, const std::string &c) {
    CubeSettings cube;
    const char *ax[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; a++) {
        cube.start[a] = req_real(v, std::string(ax[a]) + "_start", c);
        cube.end[a] = req_real(v, std::string(ax[a]) + "_end", c);
    }
    cube.value = req_real(v, "val", c);
    return cube;
}

// Settings.cpp:613-635
DiffusionSettings parse_diffusion(const xml::Element *head, const std::string &ctx) {
    const xml::Element *e = section(head, "diffusion", ctx);
    Map v = values_of(e);
    DiffusionSettings d;
    d.type = req_str(v, "type", ctx + "/diffusion");
    if (d.type == "Jacobi" || d.type == "ColoredGaussSeidel") {
        d.w = req_real(v, "w", ctx);
        d.tol_res = req_real(v, "tol_res", ctx);
        d.max_iter = req_int(v, "max_iter", ctx);
    } else if (d.type != "Explicit") {
        throw config_error("diffusion solver '" + d.type + "' has no parsing implementation.");
    }
    return d;
}

}  // namespace

Settings parse_settings(const std::string &xml_text) {
    std::unique_ptr<xml::Element> root;
    try {
        root = xml::parse(xml_text);
    } catch (const xml::parse_error &e) {
        throw config_error(e.what());
    }
    Settings s;
    const xml::Element *solver = section(root.get(), "solver", "");
    Map sv = values_of(solver);
    s.solver = req_str(sv, "description", "solver");
    const std::string &d = s.solver;
    const bool ns = d.find("NS") != std::string::npos;

    // physical_parameters, Settings.cpp:535-570
    Map pv = values_of(section(root.get(), "physical_parameters", ""));
    s.t_end = req_real(pv, "t_end", "physical_parameters");
    s.dt = req_real(pv, "dt", "physical_parameters");
    if (d != "AdvectionSolver" && d != "PressureSolver") {
        s.nu = req_real(pv, "nu", "physical_parameters");
        s.has_nu = true;
    }
    if (ns) s.rhoa = opt_real(pv, "rhoa", 1);
    s.beta = opt_real(pv, "beta", 3.34e-3);
    s.g = opt_real(pv, "g", -9.81);
    if (d == "NSTempSolver" || d == "NSTempTurbSolver" || d == "NSTempTurbConSolver" || d == "NSTempConSolver") {
        s.kappa = req_real(pv, "kappa", "physical_parameters");
        s.has_kappa = true;
    }
    if (d == "NSTempTurbConSolver" || d == "NSTempConSolver") s.gamma = req_real(pv, "gamma", "physical_parameters");

    // domain_parameters, Settings.cpp:126-148
    Map dv = values_of(section(root.get(), "domain_parameters", ""));
    if (opt_bool(dv, "enable_computational_domain", false))
        throw config_error("a computational domain smaller than the physical domain is not supported");
    s.X[0] = req_real(dv, "X1", "domain_parameters");
    s.X[1] = req_real(dv, "X2", "domain_parameters");
    s.Y[0] = req_real(dv, "Y1", "domain_parameters");
    s.Y[1] = req_real(dv, "Y2", "domain_parameters");
    s.Z[0] = req_real(dv, "Z1", "domain_parameters");
    s.Z[1] = req_real(dv, "Z2", "domain_parameters");
    s.n[0] = req_int(dv, "nx", "domain_parameters");
    s.n[1] = req_int(dv, "ny", "domain_parameters");
    s.n[2] = req_int(dv, "nz", "domain_parameters");
    for (int a = 0; a < 3; a++)
        if (s.n[a] < 1) throw config_error("domain needs at least one inner cell per axis");

    // solver sub-sections, Settings.cpp:905-970
    const bool wants_adv = d != "DiffusionSolver" && d != "DiffusionTurbSolver" && d != "PressureSolver";
    const bool wants_dif = d != "AdvectionSolver" && d != "PressureSolver";
    if (wants_adv) {
        Map av = values_of(section(solver, "advection", "solver"));
        s.advection_type = req_str(av, "type", "solver/advection");
        s.has_advection = true;
        if (s.advection_type != "SemiLagrangian")
            throw config_error("advection solver '" + s.advection_type + "' has no parsing implementation.");
    }
    if (wants_dif) {
        s.diffusion = parse_diffusion(solver, "solver");
        s.has_diffusion = true;
    }
    if (d == "DiffusionTurbSolver" || d == "NSTurbSolver" || d == "NSTempTurbSolver" || d == "NSTempTurbConSolver") {
        Map tv = values_of(section(solver, "turbulence", "solver"));
        s.turbulence_type = req_str(tv, "type", "solver/turbulence");
        s.has_turbulence = true;
        if (s.turbulence_type == "ConstSmagorinsky") s.cs = req_real(tv, "Cs", "solver/turbulence");
        else if (s.turbulence_type != "DynamicSmagorinsky")
            throw config_error("turbulence solver '" + s.turbulence_type + "' has no parsing implementation.");
    }
    if (ns) {
        Map srv = values_of(section(solver, "source", "solver"));
        s.has_source = true;
        s.source_type = req_str(srv, "type", "solver/source");
        s.force_fct = req_str(srv, "force_fct", "solver/source");
        if (s.force_fct == "Buoyancy") {
            s.use_init_values = req_bool(srv, "use_init_values", "solver/source");
            if (!s.use_init_values)
                s.ambient_temperature = req_real(srv, "ambient_temperature_value", "solver/source");
        } else if (s.force_fct == "Uniform") {  // Settings.cpp:687-696
            s.uniform_force[0] = req_real(srv, "val_x", "solver/source/Uniform");
            s.uniform_force[1] = req_real(srv, "val_y", "solver/source/Uniform");
            s.uniform_force[2] = req_real(srv, "val_z", "solver/source/Uniform");
        } else if (s.force_fct != "Zero") {
            throw config_error("source type '" + s.source_type + "' has no parsing implementation.");
        }
        s.force_dir = axes(req_str(srv, "dir", "solver/source"));
    }
    if (ns || d == "PressureSolver") {
        const xml::Element *pe = section(solver, "pressure", "solver");
        Map prv = values_of(pe);
        if (req_str(prv, "type", "solver/pressure") != "VCycleMG")
            throw config_error("pressure solver has no parsing implementation.");
        s.has_pressure = true;
        s.pressure.n_level = req_int(prv, "n_level", "solver/pressure");
        s.pressure.n_cycle = req_int(prv, "n_cycle", "solver/pressure");
        s.pressure.max_cycle = req_int(prv, "max_cycle", "solver/pressure");
        s.pressure.n_relax = req_int(prv, "n_relax", "solver/pressure");
        s.pressure.tol_res = req_real(prv, "tol_res", "solver/pressure");
        s.pressure.smoother = parse_diffusion(pe, "solver/pressure");
        s.levels = s.pressure.n_level;
    }
    if (d == "NSTempSolver" || d == "NSTempTurbSolver" || d == "NSTempConSolver" || d == "NSTempTurbConSolver") {
        const xml::Element *te = section(solver, "temperature", "solver");
        s.has_temperature = true;
        Map tav = values_of(section(te, "advection", "solver/temperature"));
        s.temp_advection_type = req_str(tav, "type", "solver/temperature/advection");
        s.temp_diffusion = parse_diffusion(te, "solver/temperature");
        if (d == "NSTempTurbSolver" || d == "NSTempTurbConSolver") {
            Map ttv = values_of(section(te, "turbulence", "solver/temperature"));
            s.temp_turbulence = req_bool(ttv, "include", "solver/temperature/turbulence");
            if (s.temp_turbulence) s.prandtl = req_real(ttv, "Pr_T", "solver/temperature/turbulence");
        }
        Map tsv = values_of(section(te, "source", "solver/temperature"));
        s.temp_source_type = req_str(tsv, "type", "solver/temperature/source");
        s.temp_dir = axes(req_str(tsv, "dir", "solver/temperature/source"));
        s.dissipation = req_bool(tsv, "dissipation", "solver/temperature/source");
        s.temp_random = req_bool(tsv, "random", "solver/temperature/source");
        s.temp_fct = req_str(tsv, "temp_fct", "solver/temperature/source");
        if (s.temp_fct == "Gauss") {
            s.gauss = parse_gauss(tsv, "solver/temperature/source/Gauss");
        } else if (s.temp_fct != "Zero" && s.temp_fct != "Buoyancy") {  // Buoyancy: BuoyancyMMS, no parameters
            throw config_error("temperature source function '" + s.temp_fct + "' has no implementation here.");
        }
    }

    // concentration, Settings.cpp:840-888 (with turbulence for NSTempTurbConSolver, :938, :953)
    if (d == "NSTempConSolver" || d == "NSTempTurbConSolver") {
        const std::string ctx = "solver/concentration";
        const xml::Element *ce = section(solver, "concentration", "solver");
        ConcentrationSettings &c = s.concentration;
        s.has_concentration = true;
        Map cav = values_of(section(ce, "advection", ctx));
        c.advection_type = req_str(cav, "type", ctx + "/advection");
        c.diffusion = parse_diffusion(ce, ctx);
        if (d == "NSTempTurbConSolver") {
            Map ctv = values_of(section(ce, "turbulence", ctx));
            c.turbulence = req_bool(ctv, "include", ctx + "/turbulence");
            if (c.turbulence) c.schmidt = req_real(ctv, "Sc_T", ctx + "/turbulence");
        }
        Map csv = values_of(section(ce, "source", ctx));
        c.source_type = req_str(csv, "type", "source");
        c.dir = axes(req_str(csv, "dir", "source"));
        if (req_bool(csv, "random", "source")) throw config_error("random concentration source is not supported");
        c.con_fct = req_str(csv, "con_fct", "source");
        if (c.con_fct == "Gauss")
            c.gauss = parse_gauss(csv, "source/Gauss");
        else if (c.con_fct == "Cube")
            c.cube = parse_cube(csv, "source/Cube");
        else
            throw config_error("concentration source function '" + c.con_fct + "' has no parsing implementation.");
    }

    // boundaries, Settings.cpp:391-402
    const xml::Element *be = section(root.get(), "boundaries", "");
    for (const auto &c : be->children)
        if (c->name == "boundary") s.boundaries.push_back(parse_boundary(c.get(), "boundaries"));

    // obstacles, Settings.cpp:404-447
    const xml::Element *oe = root->child("obstacles");
    if (oe) {
        Map ov = values_of(oe);
        if (req_bool(ov, "enabled", "obstacles")) {
            for (const auto &c : oe->children) {
                ObstacleSpec o;
                Map onv = values_of(c.get());
                o.name = req_str(onv, "name", "obstacles/obstacle");
                Map gv = values_of(section(c.get(), "geometry", "obstacles/obstacle"));
                const char *keys[6] = {"ox1", "ox2", "oy1", "oy2", "oz1", "oz2"};
                for (int k = 0; k < 6; k++) o.box[k] = req_real(gv, keys[k], "obstacles/obstacle/geometry");
                for (const auto &b : c->children)
                    if (b->name == "boundary") o.boundaries.push_back(parse_boundary(b.get(), "obstacle"));
                s.obstacles.push_back(o);
            }
        }
    }
    const xml::Element *se = root->child("surfaces");
    if (se && opt_bool(values_of(se), "enabled", false))
        throw config_error("surfaces are not supported (their reference list builder never terminates)");

    // initial conditions, Settings.cpp:330-370
    Map iv = values_of(section(root.get(), "initial_conditions", ""));
    s.ic = req_str(iv, "usr_fct", "initial_conditions");
    s.ic_random = req_bool(iv, "random", "initial_conditions");
    if (s.ic_random)
        s.ic_noise = random_settings(section(section(root.get(), "initial_conditions", ""), "random",
                                             "initial_conditions"),
                                     "initial_conditions/random");
    s.ic_values = iv;
    return s;
}

}  // namespace artss
