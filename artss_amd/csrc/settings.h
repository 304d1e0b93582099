// ARTSS input settings: the subset of the reference's Settings structs
// (utility/settings/Settings.h:37-371) that the solver hot path reads, filled
// from the same XML files with the same element/attribute names, the same
// defaults (beta 3.34e-3 and g -9.81, Settings.cpp:555-556), numbers through
// strtod/strtol like std::stod/std::stoi, and booleans as the literal "Yes".
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace artss {

// utility/Mapping.h:39-59
enum FieldType { RHO = 0, U = 1, V = 2, W = 3, P = 4, T = 5, NU = 6, NUMBER_OF_FIELD_TYPES = 7 };
enum Patch { LEFT = 0, RIGHT = 1, BOTTOM = 2, TOP = 3, FRONT = 4, BACK = 5 };
enum BoundaryCondition { BC_UNSET = -1, NEUMANN = 0, DIRICHLET = 1, PERIODIC = 2 };

class config_error : public std::runtime_error {
  public:
    explicit config_error(const std::string &m) : std::runtime_error(m) {}
};

struct BoundarySpec {
    std::vector<int> fields;
    std::vector<int> patches;
    int bc = BC_UNSET;
    double value = 0.0;
};

struct ObstacleSpec {
    std::string name;
    double box[6] = {0, 0, 0, 0, 0, 0};  // ox1 ox2 oy1 oy2 oz1 oz2
    std::vector<BoundarySpec> boundaries;
};

struct DiffusionSettings {
    std::string type;  // "Jacobi" | "ColoredGaussSeidel" | "Explicit"
    int max_iter = 0;
    double tol_res = 0;
    double w = 1;
};

struct MultigridSettings {
    int n_level = 0, n_cycle = 0, max_cycle = 0, n_relax = 0;
    double tol_res = 0;
    DiffusionSettings smoother;
};

struct GaussSettings {
    double hrr = 0, cp = 1, pos[3] = {0, 0, 0}, sigma[3] = {1, 1, 1}, tau = 1;
};

// source/Cube.cpp: a constant value on a box of cells (Settings.cpp:771-781)
struct CubeSettings {
    double start[3] = {0, 0, 0}, end[3] = {0, 0, 0}, value = 0;
};

// the <concentration> section of NSTempConSolver / NSTempTurbConSolver
// (Settings.cpp:840-888)
struct ConcentrationSettings {
    std::string advection_type;
    DiffusionSettings diffusion;  // parsed, unused: the solvers diffuse with the velocity's settings
    bool turbulence = false;
    double schmidt = 0;  // Sc_T
    std::string source_type, con_fct;  // con_fct "Gauss" or "Cube"
    std::vector<int> dir;
    GaussSettings gauss;
    CubeSettings cube;
};

// noise parameters (Settings.cpp:166-186 parse_random_parameters)
struct RandomSettings {
    bool absolute = false;     // added (Yes) or relative, x (1 + noise) (No)
    bool custom_seed = false;  // else a nondeterministic seed
    unsigned long long seed = 0;
    double step_size = 1.0;
    double range = 0.0;
};

struct Settings {
    std::string solver;  // solver description, e.g. "NSTempTurbSolver"
    double t_end = 0, dt = 0, nu = 0, kappa = 0, beta = 3.34e-3, g = -9.81, rhoa = 1, gamma = 0;
    bool has_nu = false, has_kappa = false;
    double X[2] = {0, 1}, Y[2] = {0, 1}, Z[2] = {0, 1};
    int n[3] = {1, 1, 1};

    bool has_advection = false;
    std::string advection_type;
    bool has_diffusion = false;
    DiffusionSettings diffusion;
    bool has_turbulence = false;
    std::string turbulence_type;
    double cs = 0;
    bool has_source = false;
    std::string source_type, force_fct;
    std::vector<int> force_dir;
    bool use_init_values = false;
    double ambient_temperature = 0;
    double uniform_force[3] = {0, 0, 0};  // force_fct Uniform: val_x, val_y, val_z
    bool has_pressure = false;
    MultigridSettings pressure;

    bool has_temperature = false;
    std::string temp_advection_type;
    DiffusionSettings temp_diffusion;
    bool temp_turbulence = false;
    double prandtl = 0;
    std::string temp_source_type, temp_fct;
    std::vector<int> temp_dir;
    bool dissipation = false, temp_random = false;
    GaussSettings gauss;

    bool has_concentration = false;
    ConcentrationSettings concentration;

    std::vector<BoundarySpec> boundaries;
    std::vector<ObstacleSpec> obstacles;

    std::string ic;  // initial_conditions usr_fct
    std::map<std::string, std::string> ic_values;
    bool ic_random = false;
    RandomSettings ic_noise;  // with ic_random

    int levels = 0;  // multigrid levels of the domain (main.cpp:64-68)
};

Settings parse_settings(const std::string &xml_text);

}  // namespace artss
