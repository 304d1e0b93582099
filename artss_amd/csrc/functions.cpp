#include "functions.h"

#include <cmath>
#include <cstdlib>
#include <random>

#include "settings.h"

namespace artss {

namespace {

double real_of(const Settings &s, const std::string &key) {
    auto it = s.ic_values.find(key);
    if (it == s.ic_values.end())
        throw config_error("initial condition '" + s.ic + "' needs <" + key + ">");
    char *end = nullptr;
    const double v = std::strtod(it->second.c_str(), &end);
    if (end == it->second.c_str()) throw config_error("initial condition parameter <" + key + "> is not a number");
    return v;
}

int axis_of(const Settings &s, const std::string &key) {
    auto it = s.ic_values.find(key);
    if (it == s.ic_values.end()) throw config_error("initial condition '" + s.ic + "' needs <" + key + ">");
    // Mapping::match_axis (Mapping.cpp:17-23): upper-cased, not trimmed
    const std::string &a = it->second;
    if (a == "x" || a == "X") return 0;
    if (a == "y" || a == "Y") return 1;
    if (a == "z" || a == "Z") return 2;
    throw config_error("unknown axis '" + a + "'");
}

// the geometry every function reads (DomainData getters, computational
// domain = physical domain)
struct Grid {
    size_t Nx, Ny, Nz, n;
    double X1, Y1, Z1, X2, Y2, Z2, dx, dy, dz;
    std::vector<uint64_t> inner;
    // getCoordinateK/J/I (GlobalMacrosTypes.h:71-73) and xi/yj/zk (:76-78)
    void coords(uint64_t idx, size_t &i, size_t &j, size_t &k) const {
        k = idx / (Nx * Ny);
        j = (idx - k * Nx * Ny) / Nx;
        i = idx - k * Nx * Ny - j * Nx;
    }
    double xi(size_t i) const { return X1 + (static_cast<double>(i) - 0.5) * dx; }
    double yj(size_t j) const { return Y1 + (static_cast<double>(j) - 0.5) * dy; }
    double zk(size_t k) const { return Z1 + (static_cast<double>(k) - 0.5) * dz; }
};

Grid grid_of(const Domain &d) {
    const Settings &s = d.settings();
    const LevelDesc &L = d.level(0);
    Grid g;
    g.Nx = L.N[0];
    g.Ny = L.N[1];
    g.Nz = L.N[2];
    g.n = static_cast<size_t>(L.size);
    g.X1 = s.X[0];
    g.Y1 = s.Y[0];
    g.Z1 = s.Z[0];
    g.X2 = s.X[1];
    g.Y2 = s.Y[1];
    g.Z2 = s.Z[1];
    g.dx = L.h[0];
    g.dy = L.h[1];
    g.dz = L.h[2];
    g.inner = d.inner_list(0);
    return g;
}

// Functions::random (Functions.cpp:738-770) with UniformRandom
// (randomField/UniformRandom.{h,cpp}): one draw per cell of the whole array
// in index order, an integer uniform on [-steps, steps] with steps =
// int(range / step_size), times step_size; the inner fluid cells get it
// added (absolute) or are scaled by 1 + noise (relative).  The draws come
// from the standard library's mt19937 and uniform_int_distribution, as the
// reference's do, so a custom seed reproduces the reference's noise.
void add_noise(std::vector<double> &f, const Grid &g, const RandomSettings &r) {
    const int steps = static_cast<int>(r.range / r.step_size);
    std::mt19937 mt(r.custom_seed ? static_cast<std::mt19937::result_type>(r.seed) : std::random_device()());
    std::uniform_int_distribution<int> dist(-steps, steps);
    std::vector<double> noise(g.n);
    for (size_t i = 0; i < g.n; i++) noise[i] = dist(mt) * r.step_size;
    for (uint64_t idx : g.inner) {
        if (r.absolute) f[idx] += noise[idx];
        else f[idx] *= (1 + noise[idx]);
    }
}

// Functions.cpp:471-514 at t = 0
std::vector<double> gauss_bubble(const Grid &g, const Settings &s) {
    const double u_lin = real_of(s, "u_lin"), v_lin = real_of(s, "v_lin"), w_lin = real_of(s, "w_lin");
    const double x_shift = real_of(s, "x_shift"), y_shift = real_of(s, "y_shift"), z_shift = real_of(s, "z_shift");
    const double l = real_of(s, "l");
    const double t = 0.;
    std::vector<double> out(g.n, 0.0);
    for (uint64_t idx : g.inner) {
        size_t i, j, k;
        g.coords(idx, i, j, k);
        double x_shift2 = ((g.xi(i) - x_shift) / u_lin - t) * ((g.xi(i) - x_shift) / u_lin - t);
        double y_shift2 = ((g.yj(j) - y_shift) / v_lin - t) * ((g.yj(j) - y_shift) / v_lin - t);
        double z_shift2 = 0;
        if (w_lin != 0) z_shift2 = ((g.zk(k) - z_shift) / w_lin - t) * ((g.zk(k) - z_shift) / w_lin - t);
        double quot = 1. / (2. * l * l);
        out[idx] = std::exp(-(x_shift2 + y_shift2 + z_shift2) * quot);
    }
    return out;
}

// Functions.cpp:324-358 at t = 0
std::vector<double> exp_sinus_prod(const Grid &g, const Settings &s) {
    const double nu = s.nu, A = 1.0, l = real_of(s, "l");
    const double kpinu = 3 * l * l * M_PI * M_PI * nu;
    const double t = 0.;
    std::vector<double> out(g.n, 0.0);
    for (uint64_t idx : g.inner) {
        size_t i, j, k;
        g.coords(idx, i, j, k);
        out[idx] = A * std::exp(-kpinu * t) * std::sin(l * M_PI * g.xi(i)) * std::sin(l * M_PI * g.yj(j)) *
                   std::sin(l * M_PI * g.zk(k));
    }
    return out;
}

// Functions.cpp:602-650
std::vector<double> hat(const Grid &g, const Settings &s) {
    const double start_x = real_of(s, "x1"), start_y = real_of(s, "y1"), start_z = real_of(s, "z1");
    const double end_x = real_of(s, "x2"), end_y = real_of(s, "y2"), end_z = real_of(s, "z2");
    const double val_in = real_of(s, "val_in"), val_out = real_of(s, "val_out");
    std::vector<double> out(g.n, 0.0);
    for (uint64_t idx : g.inner) {
        size_t i, j, k;
        g.coords(idx, i, j, k);
        if ((start_x <= g.xi(i) && g.xi(i) <= end_x) && (start_y <= g.yj(j) && g.yj(j) <= end_y) &&
            (start_z <= g.zk(k) && g.zk(k) <= end_z))
            out[idx] = val_in;
        else
            out[idx] = val_out;
    }
    return out;
}

// Functions.cpp:521-596: layers along `dir`, later layers winning on a border
std::vector<double> layers(const Grid &g, const Settings &s) {
    const double nl = real_of(s, "n_layers");
    if (!(nl >= 1)) throw config_error("Numbers of layers has to be at least 1");
    const int n = static_cast<int>(nl);
    const int dir = axis_of(s, "dir");
    std::vector<double> values, borders(static_cast<size_t>(n) + 1);
    for (int i = 1; i < n; i++) {
        values.push_back(real_of(s, "value_" + std::to_string(i)));
        borders[static_cast<size_t>(i)] = real_of(s, "border_" + std::to_string(i));
    }
    values.push_back(real_of(s, "value_" + std::to_string(n)));
    const double start[3] = {g.X1, g.Y1, g.Z1}, end[3] = {g.X2, g.Y2, g.Z2};
    borders[0] = start[dir];
    borders[static_cast<size_t>(n)] = end[dir];
    std::vector<double> out(g.n, 0.0);
    for (int l = 0; l < n; ++l)
        for (uint64_t idx : g.inner) {
            size_t i, j, k;
            g.coords(idx, i, j, k);
            const double c = dir == 0 ? g.xi(i) - 0.5 * g.dx : (dir == 1 ? g.yj(j) - 0.5 * g.dy : g.zk(k) - 0.5 * g.dz);
            if (borders[static_cast<size_t>(l)] <= c && c <= borders[static_cast<size_t>(l) + 1])
                out[idx] = values[static_cast<size_t>(l)];
        }
    return out;
}

}  // namespace

InitialFields initial_fields(const Domain &d) {
    const Settings &s = d.settings();
    const std::string &f = s.ic;
    InitialFields r;
    if (f == "Zero") return r;
    const Grid g = grid_of(d);
    if (f == "TaylorGreen") {
        // no reference counterpart: u = sin2πx cos2πy cos2πz, v = -cos2πx sin2πy cos2πz, w = 0
        std::vector<double> u(g.n, 0.0), v(g.n, 0.0), w(g.n, 0.0);
        for (uint64_t idx : g.inner) {
            size_t i, j, k;
            g.coords(idx, i, j, k);
            const double x = g.xi(i), y = g.yj(j), z = g.zk(k);
            u[idx] = std::sin(2. * M_PI * x) * std::cos(2. * M_PI * y) * std::cos(2. * M_PI * z);
            v[idx] = -std::cos(2. * M_PI * x) * std::sin(2. * M_PI * y) * std::cos(2. * M_PI * z);
        }
        r.arrays["u"] = std::move(u);
        r.arrays["v"] = std::move(v);
        r.arrays["w"] = std::move(w);
    } else if (f == "Uniform") {  // Functions.cpp:820-832
        const double val = real_of(s, "val");
        std::vector<double> T(g.n, 0.0);
        for (uint64_t idx : g.inner) T[idx] = val;
        r.arrays["T"] = std::move(T);
    } else if (f == "SinSinSin") {  // Functions.cpp:781-811, SolverController.cpp:126-130
        const double l = real_of(s, "l");
        std::vector<double> rhs(g.n, 0.0);
        for (uint64_t idx : g.inner) {
            size_t i, j, k;
            g.coords(idx, i, j, k);
            rhs[idx] = std::sin(l * M_PI * g.xi(i)) * std::sin(l * M_PI * g.yj(j)) * std::sin(l * M_PI * g.zk(k));
        }
        r.constants["p"] = 0.;
        r.arrays["rhs"] = std::move(rhs);
    } else if (f == "GaussBubble") {  // SolverController.cpp:97-101
        std::vector<double> a = gauss_bubble(g, s);
        r.arrays["u"] = a;
        r.arrays["v"] = a;
        r.arrays["w"] = std::move(a);
    } else if (f == "Drift") {  // Functions.cpp:297-316
        const double u_lin = real_of(s, "u_lin"), v_lin = real_of(s, "v_lin"), w_lin = real_of(s, "w_lin"),
                     pa = real_of(s, "pa");
        std::vector<double> u(g.n, 0.0), v(g.n, 0.0), w(g.n, 0.0), p(g.n, 0.0);
        for (uint64_t idx : g.inner) {
            u[idx] = u_lin;
            v[idx] = v_lin;
            w[idx] = w_lin;
            p[idx] = pa;
        }
        r.arrays["u"] = std::move(u);
        r.arrays["v"] = std::move(v);
        r.arrays["w"] = std::move(w);
        r.arrays["p"] = std::move(p);
    } else if (f == "ExpSinusProd") {  // SolverController.cpp:109-114
        std::vector<double> a = exp_sinus_prod(g, s);
        r.arrays["u"] = a;
        r.arrays["v"] = a;
        r.arrays["w"] = std::move(a);
    } else if (f == "Hat") {  // SolverController.cpp:115-119
        std::vector<double> a = hat(g, s);
        r.arrays["u"] = a;
        r.arrays["v"] = a;
        r.arrays["w"] = std::move(a);
    } else if (f == "ExpSinusSum") {  // Functions.cpp:369-420 at t = 0
        const double nu = s.nu, t = 0.;
        std::vector<double> u(g.n, 0.0), v(g.n, 0.0), w(g.n, 0.0);
        const bool three_d = g.Nz > 3;
        const double dd = three_d ? 3. : 2.;
        for (uint64_t idx : g.inner) {
            size_t i, j, k;
            g.coords(idx, i, j, k);
            if (three_d) {
                u[idx] = std::exp(-dd * nu * t) * std::sin(g.xi(i) + g.yj(j) + g.zk(k));
                v[idx] = -0.5 * std::exp(-dd * nu * t) * std::sin(g.xi(i) + g.yj(j) + g.zk(k));
                w[idx] = -0.5 * std::exp(-dd * nu * t) * std::sin(g.xi(i) + g.yj(j) + g.zk(k));
            } else {
                u[idx] = std::exp(-dd * nu * t) * std::sin(g.xi(i) + g.yj(j));
                v[idx] = -std::exp(-dd * nu * t) * std::sin(g.xi(i) + g.yj(j));
                w[idx] = 0.;
            }
        }
        r.arrays["u"] = std::move(u);
        r.arrays["v"] = std::move(v);
        r.arrays["w"] = std::move(w);
    } else if (f == "McDermott") {  // Functions.cpp:696-730 at t = 0; p then set to 0
        const double A = real_of(s, "A"), nu = s.nu, t = 0.;
        std::vector<double> u(g.n, 0.0), v(g.n, 0.0), w(g.n, 0.0);
        for (uint64_t idx : g.inner) {
            size_t i, j, k;
            g.coords(idx, i, j, k);
            u[idx] = 1. - A * std::cos(g.xi(i) - t) * std::sin(g.yj(j) - t) * std::exp(-2 * nu * t);
            v[idx] = 1. + A * std::sin(g.xi(i) - t) * std::cos(g.yj(j) - t) * std::exp(-2 * nu * t);
        }
        r.arrays["u"] = std::move(u);
        r.arrays["v"] = std::move(v);
        r.arrays["w"] = std::move(w);
        r.constants["p"] = 0.;
    } else if (f == "Vortex") {  // Functions.cpp:842-888; p then set to 0
        // the reference reads v_lin from the z entry and w_lin from the y entry
        const double u_lin = real_of(s, "u_lin"), v_lin = real_of(s, "w_lin");
        (void)real_of(s, "v_lin");
        (void)real_of(s, "pa");
        (void)real_of(s, "rhoa");
        const double L = std::fabs(g.X2 - g.X1);
        const double R_c = L / 20.;
        const double G = 0.04 * u_lin * R_c * std::sqrt(std::exp(1));
        const double GrR_c = G / (R_c * R_c);
        const double rR_c = 1. / (2. * R_c * R_c);
        std::vector<double> u(g.n, 0.0), v(g.n, 0.0), w(g.n, 0.0);
        for (uint64_t idx : g.inner) {
            size_t i, j, k;
            g.coords(idx, i, j, k);
            u[idx] = u_lin - GrR_c * g.yj(j) * std::exp(-rR_c * (g.xi(i) * g.xi(i) + g.yj(j) * g.yj(j)));
            v[idx] = v_lin + GrR_c * g.xi(i) * std::exp(-rR_c * (g.xi(i) * g.xi(i) + g.yj(j) * g.yj(j)));
        }
        r.arrays["u"] = std::move(u);
        r.arrays["v"] = std::move(v);
        r.arrays["w"] = std::move(w);
        r.constants["p"] = 0.;
    } else if (f == "LayersT") {  // SolverController.cpp:174-180
        r.arrays["T"] = layers(g, s);
    } else {
        // Beltrami, VortexY and BuoyancyMMS have no parsing implementation
        // in the reference (Settings.cpp:344-367 throws); Jet computes its end
        // indices from the domain's end coordinate (Functions.cpp:658-690),
        // which leaves its loop range undefined
        throw config_error("initial condition '" + f + "' is not supported");
    }
    // SolverController.cpp:168-180: the noise goes on T for Uniform and
    // LayersT; the reference ignores random="Yes" for the other functions
    if (s.ic_random && (f == "Uniform" || f == "LayersT")) add_noise(r.arrays["T"], g, s.ic_noise);
    return r;
}

}  // namespace artss
