#include "domain.h"
#include "residual_static.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "kernels.h"

namespace artss {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw hip_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

size_t IX(size_t i, size_t j, size_t k, size_t Nx, size_t Ny) { return i + Nx * j + Nx * Ny * k; }

// Obstacle::calculate_area_index, Obstacle.cpp:535-559
void area_index(const ObstacleLevel &o1, const ObstacleLevel &o2, size_t *c1, size_t *c2, int a, bool start) {
    if (start) {
        *c1 = o1.start[a];
        *c2 = o2.start[a];
        if (o1.start[a] > o2.start[a]) *c2 = o1.start[a] + 1;
        else if (o1.start[a] < o2.start[a]) *c1 = o2.start[a] + 1;
    } else {
        *c1 = o1.end[a];
        *c2 = o2.end[a];
        if (o1.end[a] < o2.end[a]) *c2 = o1.end[a] - 1;
        else if (o1.end[a] > o2.end[a]) *c1 = o2.end[a] - 1;
    }
}

bool overlap1(size_t a1, size_t a2, size_t b1, size_t b2) { return a1 <= b2 && a2 >= b1; }

// Face-list filter of Obstacle::circular_constraints (Obstacle.cpp:596-866).
// The reference walks both face lists with one shared loop counter and an
// enumerator of the removal area that only advances on a match; the cells it
// drops are reproduced here by the same enumeration rule.
struct FaceFilter {
    const std::vector<uint64_t> &old;
    std::vector<uint64_t> out;
    size_t pos = 0;
    int a0, a1, axis;
    size_t rs[3], re[3];
    size_t c0, c1, removing;
    bool ended = false;
    size_t Nx, Ny;
    FaceFilter(const std::vector<uint64_t> &o, int a0_, int a1_, int ax_, size_t Nx_, size_t Ny_)
        : old(o), a0(a0_), a1(a1_), axis(ax_), Nx(Nx_), Ny(Ny_) {}
    void start() {
        removing = IX(rs[0], rs[1], rs[2], Nx, Ny);
        c0 = rs[a0];
        c1 = rs[a1];
        while (pos < old.size() && old[pos] < removing) out.push_back(old[pos++]);
    }
    void next_removal(bool mark_end) {
        if (++c0 > re[a0]) {
            c0 = rs[a0];
            c1++;
            if (mark_end && c1 > re[a1]) ended = true;
        }
        size_t t[3];
        t[a0] = c0;
        t[a1] = c1;
        t[axis] = rs[axis];
        removing = IX(t[0], t[1], t[2], Nx, Ny);
    }
    void visit(bool mark_end) {
        uint64_t cur = old[pos];
        if (cur != removing) out.push_back(cur);
        else next_removal(mark_end);
    }
};

void circular_constraints(ObstacleLevel &o1, ObstacleLevel &o2, int axis, size_t Nx, size_t Ny) {
    if (o1.start[axis] - 1 != o2.end[axis]) return;
    const int a0 = axis == 0 ? 1 : 0;
    const int a1 = axis == 2 ? 1 : 2;
    if (!(overlap1(o1.start[a0], o1.end[a0], o2.start[a0], o2.end[a0]) &&
          overlap1(o1.start[a1], o1.end[a1], o2.start[a1], o2.end[a1])))
        return;
    const int p1 = 2 * axis, p2 = 2 * axis + 1;
    FaceFilter f1(o1.face[p1], a0, a1, axis, Nx, Ny), f2(o2.face[p2], a0, a1, axis, Nx, Ny);
    f1.rs[axis] = f1.re[axis] = o1.start[axis];
    f2.rs[axis] = f2.re[axis] = o2.end[axis];
    for (int a : {a0, a1}) {
        area_index(o1, o2, &f1.rs[a], &f2.rs[a], a, true);
        area_index(o1, o2, &f1.re[a], &f2.re[a], a, false);
    }
    f1.start();
    f2.start();
    for (; f1.pos < f1.old.size() && f2.pos < f2.old.size() && !f1.ended && !f2.ended; f1.pos++, f2.pos++) {
        f1.visit(true);
        f2.visit(true);
    }
    for (FaceFilter *f : {&f1, &f2}) {
        if (!f->ended)
            for (; f->pos < f->old.size() && f->c1 <= f->re[f->a1]; f->pos++) f->visit(false);
        for (; f->pos < f->old.size(); f->pos++) f->out.push_back(f->old[f->pos]);
    }
    o1.face[p1] = std::move(f1.out);
    o2.face[p2] = std::move(f2.out);
}

}  // namespace

Domain::Domain(const Settings &s, int device) : m_settings(s), m_device(device) {
    m_levels = s.levels;
    if (m_levels < 0 || m_levels > 16) throw config_error("multigrid levels out of range");
    build_levels();
    build_obstacles();
    // fluid cells per level, counted once (the pressure solve's warm start
    // asks for them every step)
    m_inner_count.resize(m_level.size());
    for (size_t l = 0; l < m_level.size(); l++) {
        const LevelDesc &L = m_level[l];
        size_t n = 0;
        for (int k = 1; k <= L.n[2]; k++)
            for (int j = 1; j <= L.n[1]; j++)
                for (int i = 1; i <= L.n[0]; i++)
                    if (!in_obstacle(L, i, j, k)) n++;
        m_inner_count[l] = n;
    }
    if (device >= 0) {
        ARTSS_HIP(hipSetDevice(device));
        ARTSS_HIP(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking));
    }
    build_boundary_tables();
    if (device < 0) return;  // host-only: geometry and index lists, no device residency
    m_partial_cap = 1;
    for (int l = 0; l <= m_levels; l++) m_partial_cap = std::max(m_partial_cap, k::partial_count(m_level[l]));
    ARTSS_HIP(hipMalloc(&m_partial, sizeof(double) * static_cast<size_t>(m_partial_cap)));
    ARTSS_HIP(hipMalloc(&m_sum_dev, sizeof(double)));
    ARTSS_HIP(hipHostMalloc(&m_sum_host, sizeof(double), hipHostMallocDefault));
    static_scratch_setup(device);
}

Domain::~Domain() {
    if (!m_stream) return;  // host-only domain
    timer_clear();
    (void)hipStreamSynchronize(m_stream);
    (void)hipFree(m_partial);
    (void)hipFree(m_sum_dev);
    (void)hipHostFree(m_sum_host);
    (void)hipFree(m_faces_dev);
    (void)hipFree(m_params_dev);
    (void)hipFree(m_params_hom_dev);
    (void)hipFree(m_bits_dev);
    (void)hipFree(m_deps_dev);
    (void)hipFree(m_rest_dev);
    (void)hipStreamDestroy(m_stream);
}

void Domain::require_device() const {
    if (!m_stream) throw std::invalid_argument("domain was created without a device (host-only)");
}

// DomainData.cpp:12-56 and the spacing/index getters of DomainData.h:144-171,
// for a computational domain equal to the physical one.
void Domain::build_levels() {
    const Settings &s = m_settings;
    const double lo[3] = {s.X[0], s.Y[0], s.Z[0]}, hi[3] = {s.X[1], s.Y[1], s.Z[1]};
    double len[3];
    for (int a = 0; a < 3; a++) len[a] = std::fabs(hi[a] - lo[a]);
    m_level.resize(m_levels + 1);
    size_t n[3] = {static_cast<size_t>(s.n[0]), static_cast<size_t>(s.n[1]), static_cast<size_t>(s.n[2])};
    for (int l = 0; l <= m_levels; l++) {
        LevelDesc &L = m_level[l];
        std::memset(&L, 0, sizeof(L));
        if (l > 0)
            for (int a = 0; a < 3; a++) n[a] = (n[a] == 1) ? 1 : static_cast<size_t>(std::round(n[a] / 2));
        L.size = 1;
        for (int a = 0; a < 3; a++) {
            L.n[a] = static_cast<int>(n[a]);
            L.h[a] = len[a] / static_cast<double>(n[a]);
            L.N[a] = static_cast<int>(static_cast<size_t>(std::round(len[a] / L.h[a]) + 2));
            L.size *= L.N[a];
        }
        if (L.N[0] != L.n[0] + 2 || L.N[1] != L.n[1] + 2 || L.N[2] != L.n[2] + 2)
            throw config_error("inconsistent level sizes");
        // kernels decode cell indices and fill items in 32-bit arithmetic
        if (L.size >= (1LL << 31)) throw config_error("grid too large: a level needs fewer than 2^31 cells");
    }
}

// Obstacle.cpp:13-191, 449-478, 562-866; Multigrid.cpp:180-265, 438-514.
void Domain::build_obstacles() {
    const auto &os = m_settings.obstacles;
    const int n_obst = static_cast<int>(os.size());
    if (n_obst > kMaxObstacles) throw config_error("too many obstacles (limit " + std::to_string(kMaxObstacles) + ")");
    m_obst.assign(m_levels + 1, std::vector<ObstacleLevel>(n_obst));
    const double lo[3] = {m_settings.X[0], m_settings.Y[0], m_settings.Z[0]};
    auto init_faces = [&](ObstacleLevel &o, int l) {
        const LevelDesc &L = m_level[l];
        const size_t Nx = L.N[0], Ny = L.N[1];
        size_t st[3];
        for (int a = 0; a < 3; a++) st[a] = o.end[a] - o.start[a] + 1;
        size_t nface[6] = {st[2] * st[1], st[2] * st[1], st[2] * st[0], st[2] * st[0], st[1] * st[0], st[1] * st[0]};
        // remove_cells_at_boundary: faces on the domain boundary have no fluid neighbour
        if (o.start[2] <= 1) nface[FRONT] = 0;
        if (o.end[2] >= static_cast<size_t>(L.n[2])) nface[BACK] = 0;
        if (o.start[1] <= 1) nface[BOTTOM] = 0;
        if (o.end[1] >= static_cast<size_t>(L.n[1])) nface[TOP] = 0;
        if (o.start[0] <= 1) nface[LEFT] = 0;
        if (o.end[0] >= static_cast<size_t>(L.n[0])) nface[RIGHT] = 0;
        for (int p = 0; p < 6; p++) o.face[p].clear();
        // face cells in the reference's list order (create_obstacle, Obstacle.cpp:104-191)
        for (size_t k = o.start[2]; k <= o.end[2]; k++)
            for (size_t j = o.start[1]; j <= o.end[1]; j++)
                for (size_t i = o.start[0]; i <= o.end[0]; i++) {
                    const uint64_t idx = IX(i, j, k, Nx, Ny);
                    if (nface[LEFT] && i == o.start[0]) o.face[LEFT].push_back(idx);
                    if (nface[RIGHT] && i == o.end[0]) o.face[RIGHT].push_back(idx);
                    if (nface[BOTTOM] && j == o.start[1]) o.face[BOTTOM].push_back(idx);
                    if (nface[TOP] && j == o.end[1]) o.face[TOP].push_back(idx);
                    if (nface[FRONT] && k == o.start[2]) o.face[FRONT].push_back(idx);
                    if (nface[BACK] && k == o.end[2]) o.face[BACK].push_back(idx);
                }
    };
    for (int id = 0; id < n_obst; id++) {
        ObstacleLevel &o = m_obst[0][id];
        for (int a = 0; a < 3; a++) {
            const double h = m_level[0].h[a];
            const int s = static_cast<int>(std::round((-lo[a] + os[id].box[2 * a]) / h));
            const int e = static_cast<int>(std::round((-lo[a] + os[id].box[2 * a + 1]) / h));
            if (s + 1 < 1 || e > m_level[0].n[a] || s + 1 > e)
                throw config_error("obstacle '" + os[id].name + "' does not lie inside the domain");
            o.start[a] = static_cast<size_t>(s + 1);
            o.end[a] = static_cast<size_t>(e);
        }
        init_faces(o, 0);
    }
    for (int a = 0; a < n_obst; a++)
        for (int b = a + 1; b < n_obst; b++) {
            bool ov = true;
            for (int x = 0; x < 3; x++)
                ov = ov && overlap1(m_obst[0][a].start[x], m_obst[0][a].end[x], m_obst[0][b].start[x],
                                    m_obst[0][b].end[x]);
            if (ov) throw config_error("obstacles '" + os[a].name + "' and '" + os[b].name + "' overlap");
        }
    // DomainController::detect_neighbouring_obstacles -> Obstacle::remove_circular_constraints
    const size_t Nx0 = m_level[0].N[0], Ny0 = m_level[0].N[1];
    for (int a = 0; a < n_obst; a++)
        for (int b = a + 1; b < n_obst; b++)
            for (int axis = 0; axis < 3; axis++) {
                ObstacleLevel &o1 = m_obst[0][a], &o2 = m_obst[0][b];
                if (o1.end[axis] + 1 == o2.start[axis]) circular_constraints(o2, o1, axis, Nx0, Ny0);
                else circular_constraints(o1, o2, axis, Nx0, Ny0);
            }
    // dominant restriction: coarse box from (start+1)/2 .. (end+1)/2, faces rebuilt
    for (int l = 1; l <= m_levels; l++)
        for (int id = 0; id < n_obst; id++) {
            ObstacleLevel &o = m_obst[l][id];
            for (int a = 0; a < 3; a++) {
                o.start[a] = (m_obst[l - 1][id].start[a] + 1) / 2;
                o.end[a] = (m_obst[l - 1][id].end[a] + 1) / 2;
            }
            init_faces(o, l);
        }
    for (int l = 0; l <= m_levels; l++) {
        LevelDesc &L = m_level[l];
        L.n_obst = n_obst;
        for (int id = 0; id < n_obst; id++)
            for (int a = 0; a < 3; a++) {
                L.obst[id].lo[a] = static_cast<int>(m_obst[l][id].start[a]);
                L.obst[id].hi[a] = static_cast<int>(m_obst[l][id].end[a]);
            }
    }
}

// BoundaryDataController + DomainBoundary/ObstacleBoundary parameters
// (DomainBoundary.cpp:59-239, ObstacleBoundary.cpp:62-213), precomputed per
// level and field type.
void Domain::build_boundary_tables() {
    const int n_obst = obstacle_count();
    struct PatchBC {
        int bc = BC_UNSET;
        double value = 0;
    };
    // BoundaryData::add_boundary_condition: later entries overwrite earlier ones
    auto collect = [](const std::vector<BoundarySpec> &specs, bool has[NUMBER_OF_FIELD_TYPES],
                      PatchBC tab[NUMBER_OF_FIELD_TYPES][6]) {
        for (const auto &b : specs)
            for (int f : b.fields) {
                has[f] = true;
                for (int p : b.patches) tab[f][p] = PatchBC{b.bc, b.value};
            }
    };
    bool dom_has[NUMBER_OF_FIELD_TYPES] = {};
    PatchBC dom[NUMBER_OF_FIELD_TYPES][6];
    collect(m_settings.boundaries, dom_has, dom);
    for (int f = 0; f < NUMBER_OF_FIELD_TYPES; f++)
        if (dom_has[f])
            for (int p = 0; p < 6; p++)
                if (dom[f][p].bc == BC_UNSET)
                    throw config_error("field type " + std::to_string(f) + " has boundary data for some patches only");
    std::vector<std::array<bool, NUMBER_OF_FIELD_TYPES>> ob_has(n_obst);
    std::vector<std::array<std::array<PatchBC, 6>, NUMBER_OF_FIELD_TYPES>> ob(n_obst);
    for (int id = 0; id < n_obst; id++) {
        bool has[NUMBER_OF_FIELD_TYPES] = {};
        PatchBC tab[NUMBER_OF_FIELD_TYPES][6];
        collect(m_settings.obstacles[id].boundaries, has, tab);
        for (int f = 0; f < NUMBER_OF_FIELD_TYPES; f++) {
            ob_has[id][f] = has[f];
            for (int p = 0; p < 6; p++) {
                ob[id][f][p] = tab[f][p];
                if (has[f] && tab[f][p].bc == BC_UNSET)
                    throw config_error("obstacle '" + m_settings.obstacles[id].name +
                                       "' has boundary data for some patches only");
            }
        }
    }

    m_dom_bc.assign(m_levels + 1, std::vector<DomainBC>(NUMBER_OF_FIELD_TYPES));
    m_waves.assign(m_levels + 1, std::vector<std::vector<Wave>>(NUMBER_OF_FIELD_TYPES));
    m_param_offset.assign(m_levels + 1, std::vector<long long>(NUMBER_OF_FIELD_TYPES, 0));
    m_simple.assign(m_levels + 1, std::vector<char>(NUMBER_OF_FIELD_TYPES, 0));
    std::vector<FaceEntry> entries;
    std::vector<FaceParam> params;
    std::vector<uint32_t> bits;
    std::vector<DepGhost> deps;
    std::vector<uint32_t> rest;
    m_idem.assign(m_levels + 1, std::vector<char>(NUMBER_OF_FIELD_TYPES, 0));
    m_face_values_zero.assign(m_levels + 1, std::vector<char>(NUMBER_OF_FIELD_TYPES, 1));
    m_view_bits.assign(m_levels + 1, std::vector<long long>(NUMBER_OF_FIELD_TYPES, 0));
    m_shell.assign(m_levels + 1, std::vector<char>(NUMBER_OF_FIELD_TYPES, 0));
    m_view_deps.assign(m_levels + 1, std::vector<long long>(NUMBER_OF_FIELD_TYPES, 0));
    m_view_ndeps.assign(m_levels + 1, std::vector<long long>(NUMBER_OF_FIELD_TYPES, 0));
    m_view_rest.assign(m_levels + 1, std::vector<long long>(NUMBER_OF_FIELD_TYPES, 0));
    m_view_nrest.assign(m_levels + 1, std::vector<long long>(NUMBER_OF_FIELD_TYPES, 0));

    for (int l = 0; l <= m_levels; l++) {
        const LevelDesc &L = m_level[l];
        const long long Nx = L.N[0], Ny = L.N[1];
        for (int f = 0; f < NUMBER_OF_FIELD_TYPES; f++) {
            // ---- domain patches
            DomainBC &d = m_dom_bc[l][f];
            std::memset(&d, 0, sizeof(d));
            d.active = dom_has[f] ? 1 : 0;
            for (int p = 0; p < 6 && d.active; p++) {
                const int axis = p / 2;
                const bool high = (p % 2) == 1;
                const PatchBC &b = dom[f][p];
                const double v = (l == 0) ? b.value : 0.0;
                if (b.bc == DIRICHLET) {
                    d.sign[p] = -1.0;
                    d.value[p] = v * 2;
                    d.target[p] = high ? L.N[axis] - 2 : 1;
                } else if (b.bc == NEUMANN) {
                    // dz for FRONT/BACK, dy for BOTTOM/TOP and, as in the reference, dz for LEFT/RIGHT
                    const double h = (p == BOTTOM || p == TOP) ? L.h[1] : L.h[2];
                    d.sign[p] = 1.0;
                    d.value[p] = v * h;
                    d.target[p] = high ? L.N[axis] - 2 : 1;
                } else {  // periodic: reference cell n cells away, DomainBoundary.cpp:147-182
                    d.sign[p] = 1.0;
                    d.value[p] = 0.0;
                    d.target[p] = high ? L.N[axis] - 1 - L.n[axis] : L.n[axis];
                }
            }
            // ---- obstacle faces: parameter table + write schedule
            m_param_offset[l][f] = static_cast<long long>(params.size());
            m_face_values_zero[l][f] = 1;
            for (int id = 0; id < n_obst; id++)
                for (int p = 0; p < 6; p++) {
                    FaceParam q{};
                    const PatchBC &b = ob[id][f][p];
                    if (ob_has[id][f] && (b.bc == DIRICHLET || b.bc == NEUMANN)) {
                        const double v = (l > 0) ? 0.0 : b.value;
                        q.active = 1;
                        if (b.bc == DIRICHLET) {
                            q.sign = -1.0;
                            q.value = v * 2;
                        } else {
                            const double h = (p == FRONT || p == BACK) ? L.h[2] : (p == BOTTOM || p == TOP) ? L.h[1] : L.h[0];
                            q.sign = 1.0;
                            q.value = -(v * h);
                        }
                    }
                    if (q.active && q.value != 0.0) m_face_values_zero[l][f] = 0;
                    params.push_back(q);
                }
            // writes in reference order (Multigrid.cpp:559-562, ObstacleBoundary.cpp:178-209)
            struct Write {
                uint64_t cell, read;
                int32_t ref, param;
                bool live;
                int wave;
            };
            std::vector<Write> ws;
            for (int id = 0; id < n_obst; id++) {
                if (!ob_has[id][f]) continue;
                for (int p = 0; p < 6; p++) {
                    const FaceParam &q = params[m_param_offset[l][f] + id * 6 + p];
                    if (!q.active) continue;
                    const long long step = (p == FRONT || p == BACK) ? Nx * Ny : (p == BOTTOM || p == TOP) ? Nx : 1;
                    const long long ref = (p == FRONT || p == BOTTOM || p == LEFT) ? -step : step;
                    for (uint64_t c : m_obst[l][id].face[p])
                        ws.push_back(Write{c, static_cast<uint64_t>(static_cast<long long>(c) + ref),
                                           static_cast<int32_t>(ref), id * 6 + p, true, 0});
                }
            }
            // drop writes overwritten before anyone reads them
            std::unordered_map<uint64_t, int> next;  // 1: next event is a write, 2: a read
            for (size_t t = ws.size(); t-- > 0;) {
                auto it = next.find(ws[t].cell);
                if (it != next.end() && it->second == 1) {
                    ws[t].live = false;
                    continue;
                }
                next[ws[t].cell] = 1;
                next[ws[t].read] = 2;
            }
            // waves: a write runs after the writes of the cell it reads and after
            // every read of the cell it overwrites
            std::unordered_map<uint64_t, int> last_write, last_read;
            int n_waves = 0;
            for (auto &w : ws) {
                if (!w.live) continue;
                int wave = 0;
                auto a = last_write.find(w.read);
                if (a != last_write.end()) wave = std::max(wave, a->second + 1);
                auto b = last_read.find(w.cell);
                if (b != last_read.end()) wave = std::max(wave, b->second + 1);
                auto c = last_write.find(w.cell);
                if (c != last_write.end()) wave = std::max(wave, c->second + 1);
                w.wave = wave;
                last_write[w.cell] = wave;
                auto &r = last_read[w.read];
                r = std::max(r, wave);
                n_waves = std::max(n_waves, wave + 1);
            }
            // Classify the fill.  W = cells it writes (live faces, ghosts).
            //  idempotent: no face write reads a cell of W, so one wave of
            //    faces suffices and the result depends on non-W cells only;
            //  simple: idempotent and no ghost reads a face cell either, so
            //    faces and ghosts are independent.
            std::vector<char> written(static_cast<size_t>(L.size), 0);
            std::vector<long long> ghost_src;  // per ghost (k_bc_domain order) the cell its chain reads
            std::vector<long long> ghost_cell; // and the ghost cell itself
            {
                for (const auto &w : ws)
                    if (w.live) written[w.cell] = 2;  // 2: face-written
                const long long Nz = L.N[2];
                if (d.active) {
                    const long long fx = Ny * Nz, fy = (Nx - 2) * Nz, fz = (Nx - 2) * (Ny - 2);
                    const long long n_ghost = 2 * fx + 2 * fy + 2 * fz;
                    ghost_src.resize(static_cast<size_t>(n_ghost));
                    ghost_cell.resize(static_cast<size_t>(n_ghost));
                    for (long long t0 = 0; t0 < n_ghost; t0++) {
                        long long t = t0, i, j, k;
                        if (t < 2 * fx) {
                            i = (t / fx) ? Nx - 1 : 0;
                            j = (t % fx) % Ny;
                            k = (t % fx) / Ny;
                        } else if ((t -= 2 * fx) < 2 * fy) {
                            j = (t / fy) ? Ny - 1 : 0;
                            i = 1 + (t % fy) % (Nx - 2);
                            k = (t % fy) / (Nx - 2);
                        } else {
                            t -= 2 * fy;
                            k = (t / fz) ? Nz - 1 : 0;
                            i = 1 + (t % fz) % (Nx - 2);
                            j = 1 + (t % fz) / (Nx - 2);
                        }
                        written[i + Nx * (j + Ny * k)] = 1;
                        const long long ii = (i == 0 || i == Nx - 1) ? d.target[i == 0 ? 0 : 1] : i;
                        const long long jj = (j == 0 || j == Ny - 1) ? d.target[j == 0 ? 2 : 3] : j;
                        const long long kk = (k == 0 || k == Nz - 1) ? d.target[k == 0 ? 4 : 5] : k;
                        ghost_src[static_cast<size_t>(t0)] = ii + Nx * (jj + Ny * kk);
                        ghost_cell[static_cast<size_t>(t0)] = i + Nx * (j + Ny * k);
                    }
                }
                bool idem = true;
                for (const auto &w : ws)
                    if (w.live && written[w.read]) idem = false;
                bool simple = idem;
                for (long long src : ghost_src)
                    if (written[static_cast<size_t>(src)]) simple = false;
                m_simple[l][f] = simple ? 1 : 0;
                m_idem[l][f] = idem ? 1 : 0;
            }
            for (int wv = 0; wv < n_waves; wv++) {
                Wave W{static_cast<long long>(entries.size()), 0};
                for (const auto &w : ws)
                    if (w.live && w.wave == wv) {
                        if (w.cell > 0xffffffffULL) throw config_error("grid too large for 32-bit face indices");
                        entries.push_back(FaceEntry{static_cast<uint32_t>(w.cell), w.ref,
                                                    w.param, 0});
                        W.count++;
                    }
                m_waves[l][f].push_back(W);
            }
            // device tables of an idempotent fill: bitmap of W and the ghosts
            // whose chain starts on a face cell, with that face's entry
            m_view_bits[l][f] = static_cast<long long>(bits.size());
            m_view_deps[l][f] = static_cast<long long>(deps.size());
            m_view_ndeps[l][f] = 0;
            if (m_idem[l][f]) {
                const size_t words = (static_cast<size_t>(L.size) + 31) / 32;
                const size_t b0 = bits.size();
                bits.resize(b0 + words, 0u);
                for (size_t c = 0; c < written.size(); c++)
                    if (written[c]) bits[b0 + c / 32] |= 1u << (c % 32);
                bool shell = true;
                for (long long k = 0; k < L.N[2] && shell; k++)
                    for (long long j = 0; j < L.N[1] && shell; j++)
                        for (long long i = 0; i < L.N[0]; i++) {
                            const bool ghost = i == 0 || j == 0 || k == 0 || i == L.N[0] - 1 || j == L.N[1] - 1 ||
                                               k == L.N[2] - 1;
                            if (ghost && !written[static_cast<size_t>(i + L.N[0] * (j + L.N[1] * k))]) {
                                shell = false;
                                break;
                            }
                        }
                m_shell[l][f] = shell ? 1 : 0;
                if (!m_waves[l][f].empty()) {
                    std::unordered_map<uint64_t, long long> entry_of;
                    const Wave &W0 = m_waves[l][f][0];
                    for (long long e = W0.offset; e < W0.offset + W0.count; e++) entry_of[entries[e].idx] = e - W0.offset;
                    const long long Nz = L.N[2];
                    const long long fx = Ny * Nz, fy = (Nx - 2) * Nz;
                    for (size_t t0 = 0; t0 < ghost_src.size(); t0++) {
                        const long long src = ghost_src[t0];
                        if (written[static_cast<size_t>(src)] != 2) continue;
                        long long t = static_cast<long long>(t0), i, j, k;
                        if (t < 2 * fx) {
                            i = (t / fx) ? Nx - 1 : 0;
                            j = (t % fx) % Ny;
                            k = (t % fx) / Ny;
                        } else if ((t -= 2 * fx) < 2 * fy) {
                            j = (t / fy) ? Ny - 1 : 0;
                            i = 1 + (t % fy) % (Nx - 2);
                            k = (t % fy) / (Nx - 2);
                        } else {
                            const long long fz = (Nx - 2) * (Ny - 2);
                            t -= 2 * fy;
                            k = (t / fz) ? Nz - 1 : 0;
                            i = 1 + (t % fz) % (Nx - 2);
                            j = 1 + (t % fz) / (Nx - 2);
                        }
                        deps.push_back(DepGhost{static_cast<uint32_t>(i + Nx * (j + Ny * k)),
                                                static_cast<int32_t>(entry_of.at(static_cast<uint64_t>(src)))});
                    }
                }
                m_view_ndeps[l][f] = static_cast<long long>(deps.size()) - m_view_deps[l][f];
                // ghosts the sweep kernels' main grid does not write
                // (stencil.h face_ghosts): edges, corners and face ghosts
                // whose chain starts on a cell that is not inner fluid; the
                // ghosts reading a face cell are the dependent ones above
                m_view_rest[l][f] = static_cast<long long>(rest.size());
                const long long Nz = L.N[2];
                for (size_t t0 = 0; t0 < ghost_src.size(); t0++) {
                    const long long src = ghost_src[t0], g = ghost_cell[t0];
                    if (written[static_cast<size_t>(src)]) continue;
                    const long long i = g % Nx, j = (g / Nx) % Ny, k = g / (Nx * Ny);
                    const int on = (i == 0 || i == Nx - 1) + (j == 0 || j == Ny - 1) + (k == 0 || k == Nz - 1);
                    const int si = static_cast<int>(src % Nx), sj = static_cast<int>((src / Nx) % Ny),
                              sk = static_cast<int>(src / (Nx * Ny));
                    const bool fluid = si >= 1 && si <= L.n[0] && sj >= 1 && sj <= L.n[1] && sk >= 1 &&
                                       sk <= L.n[2] && !in_obstacle(L, si, sj, sk);
                    if (on == 1 && fluid) continue;
                    if (g > 0xffffffffLL) throw config_error("grid too large for 32-bit ghost indices");
                    rest.push_back(static_cast<uint32_t>(g));
                }
                m_view_nrest[l][f] = static_cast<long long>(rest.size()) - m_view_rest[l][f];
            }
        }
    }
    m_face_entries = static_cast<long long>(entries.size());
    if (m_device < 0) return;
    if (!entries.empty()) {
        ARTSS_HIP(hipMalloc(&m_faces_dev, sizeof(FaceEntry) * entries.size()));
        ARTSS_HIP(hipMemcpy(m_faces_dev, entries.data(), sizeof(FaceEntry) * entries.size(), hipMemcpyHostToDevice));
    }
    if (!params.empty()) {
        ARTSS_HIP(hipMalloc(&m_params_dev, sizeof(FaceParam) * params.size()));
        ARTSS_HIP(hipMemcpy(m_params_dev, params.data(), sizeof(FaceParam) * params.size(), hipMemcpyHostToDevice));
        // the same table with every boundary value zero (linear part of the fill)
        std::vector<FaceParam> hom = params;
        for (FaceParam &q : hom) q.value = 0.0;
        ARTSS_HIP(hipMalloc(&m_params_hom_dev, sizeof(FaceParam) * hom.size()));
        ARTSS_HIP(hipMemcpy(m_params_hom_dev, hom.data(), sizeof(FaceParam) * hom.size(), hipMemcpyHostToDevice));
    }
    if (!bits.empty()) {
        ARTSS_HIP(hipMalloc(&m_bits_dev, sizeof(uint32_t) * bits.size()));
        ARTSS_HIP(hipMemcpy(m_bits_dev, bits.data(), sizeof(uint32_t) * bits.size(), hipMemcpyHostToDevice));
    }
    if (!deps.empty()) {
        ARTSS_HIP(hipMalloc(&m_deps_dev, sizeof(DepGhost) * deps.size()));
        ARTSS_HIP(hipMemcpy(m_deps_dev, deps.data(), sizeof(DepGhost) * deps.size(), hipMemcpyHostToDevice));
    }
    if (!rest.empty()) {
        ARTSS_HIP(hipMalloc(&m_rest_dev, sizeof(uint32_t) * rest.size()));
        ARTSS_HIP(hipMemcpy(m_rest_dev, rest.data(), sizeof(uint32_t) * rest.size(), hipMemcpyHostToDevice));
    }
}

BCView Domain::bc_view(int level, int ft, bool homogeneous) const {
    if (!m_idem[level][ft]) throw std::logic_error("bc_view of a boundary fill that is not idempotent");
    BCView v{};
    const auto &W = m_waves[level][ft];
    v.faces = W.empty() ? nullptr : m_faces_dev + W[0].offset;
    v.n_faces = W.empty() ? 0 : W[0].count;
    FaceParam *params = homogeneous ? m_params_hom_dev : m_params_dev;
    v.params = params ? params + m_param_offset[level][ft] : nullptr;
    v.deps = m_deps_dev ? m_deps_dev + m_view_deps[level][ft] : nullptr;
    v.n_deps = m_view_ndeps[level][ft];
    v.rest = m_rest_dev ? m_rest_dev + m_view_rest[level][ft] : nullptr;
    v.n_rest = m_view_nrest[level][ft];
    v.wbits = m_bits_dev + m_view_bits[level][ft];
    v.bc = m_dom_bc[level][ft];
    if (homogeneous)
        for (double &x : v.bc.value) x = 0.0;
    v.n_ghosts = v.bc.active ? ghost_cells(level) : 0;
    v.shell_filled = m_shell[level][ft];
    return v;
}

bool Domain::bc_values_zero(int level, int ft) const {
    const DomainBC &d = m_dom_bc[level][ft];
    for (int p = 0; p < 6; p++)
        if (d.active && d.value[p] != 0.0) return false;
    return m_face_values_zero[level][ft] != 0;
}

long long Domain::ghost_cells(int level) const {
    const LevelDesc &L = m_level[level];
    const long long Nx = L.N[0], Ny = L.N[1], Nz = L.N[2];
    return 2 * Ny * Nz + 2 * (Nx - 2) * Nz + 2 * (Nx - 2) * (Ny - 2);
}

void Domain::apply_boundary(double *f, int ft, int level, const int *gate) {
    require_device();
    if (ft < 0 || ft >= NUMBER_OF_FIELD_TYPES) return;  // UNKNOWN_FIELD has no boundary data
    for (const Wave &w : m_waves[level][ft])
        timed("bc_faces", level, [&] {
            k::bc_faces(m_stream, m_faces_dev + w.offset, w.count, m_params_dev + m_param_offset[level][ft], f, gate);
        });
    timed("bc_domain", level, [&] { k::bc_domain(m_stream, m_level[level], m_dom_bc[level][ft], f, gate); });
}

void Domain::apply_boundary_merged(double *f, int ft, int level) {
    require_device();
    if (ft < 0 || ft >= NUMBER_OF_FIELD_TYPES) return;
    if (!m_simple[level][ft]) throw std::logic_error("apply_boundary_merged on a fill that is not simple");
    const auto &W = m_waves[level][ft];
    const long long off = W.empty() ? 0 : W[0].offset, n = W.empty() ? 0 : W[0].count;
    timed("bc_merged", level, [&] {
        k::bc_merged(m_stream, m_level[level], m_dom_bc[level][ft], m_faces_dev + off, n,
                     m_params_dev + m_param_offset[level][ft], f);
    });
}

void Domain::apply_boundary_linear(double *f, int ft, int level) {
    require_device();
    if (ft < 0 || ft >= NUMBER_OF_FIELD_TYPES) return;
    if (m_idem[level][ft]) {
        timed("bc_merged", level, [&] { k::fill_bc(m_stream, m_level[level], bc_view(level, ft, true), f); });
        return;
    }
    for (const Wave &w : m_waves[level][ft])
        timed("bc_faces", level, [&] {
            k::bc_faces(m_stream, m_faces_dev + w.offset, w.count, m_params_hom_dev + m_param_offset[level][ft], f);
        });
    DomainBC bc = m_dom_bc[level][ft];
    for (double &x : bc.value) x = 0.0;
    timed("bc_domain", level, [&] { k::bc_domain(m_stream, m_level[level], bc, f); });
}

void Domain::apply_faces(double *f, int ft, int level) {
    require_device();
    if (ft < 0 || ft >= NUMBER_OF_FIELD_TYPES) return;
    for (const Wave &w : m_waves[level][ft])
        timed("bc_faces", level, [&] {
            k::bc_faces(m_stream, m_faces_dev + w.offset, w.count, m_params_dev + m_param_offset[level][ft], f);
        });
}

void Domain::reduce_async(const double *partial, int n) {
    require_device();
    k::reduce_sum(m_stream, partial, n, m_sum_dev);
    ARTSS_HIP(hipMemcpyAsync(m_sum_host, m_sum_dev, sizeof(double), hipMemcpyDeviceToHost, m_stream));
}

double Domain::reduce(const double *partial, int n) {
    reduce_async(partial, n);
    ARTSS_HIP(hipStreamSynchronize(m_stream));
    return *m_sum_host;
}

size_t Domain::inner_cells(int level) const { return m_inner_count.at(static_cast<size_t>(level)); }

std::vector<uint64_t> Domain::inner_list(int level) const {
    const LevelDesc &L = m_level[level];
    std::vector<uint64_t> out;
    for (int k = 1; k <= L.n[2]; k++)
        for (int j = 1; j <= L.n[1]; j++)
            for (int i = 1; i <= L.n[0]; i++)
                if (!in_obstacle(L, i, j, k)) out.push_back(IX(i, j, k, L.N[0], L.N[1]));
    return out;
}

std::vector<uint64_t> Domain::domain_list(int level) const {
    const LevelDesc &L = m_level[level];
    std::vector<uint64_t> out;
    for (int k = 0; k < L.N[2]; k++)
        for (int j = 0; j < L.N[1]; j++)
            for (int i = 0; i < L.N[0]; i++) {
                const bool ghost = i == 0 || j == 0 || k == 0 || i == L.N[0] - 1 || j == L.N[1] - 1 || k == L.N[2] - 1;
                if (ghost || !in_obstacle(L, i, j, k)) out.push_back(IX(i, j, k, L.N[0], L.N[1]));
            }
    return out;
}

std::vector<uint64_t> Domain::obstacle_list(int level) const {
    // level 0 keeps duplicates (Algorithm::merge_sort), coarser levels are
    // unique (merge_sort_with_duplicates); level-0 obstacles never overlap.
    const LevelDesc &L = m_level[level];
    std::vector<uint64_t> out;
    for (int k = 1; k <= L.n[2]; k++)
        for (int j = 1; j <= L.n[1]; j++)
            for (int i = 1; i <= L.n[0]; i++)
                if (in_obstacle(L, i, j, k)) out.push_back(IX(i, j, k, L.N[0], L.N[1]));
    return out;
}

namespace {
bool is_obstacle_cell_real(const ObstacleLevel &o, const double c[3]) {
    return static_cast<double>(o.start[0]) <= c[0] && c[0] <= static_cast<double>(o.end[0]) &&
           static_cast<double>(o.start[1]) <= c[1] && c[1] <= static_cast<double>(o.end[1]) &&
           static_cast<double>(o.start[2]) <= c[2] && c[2] <= static_cast<double>(o.end[2]);
}
// Obstacle::intersection, Obstacle.cpp:909-978: is the segment between the
// cell centres of `src` and `cell` cut by the obstacle?
bool blocks_sight(const ObstacleLevel &o, const size_t src[3], const size_t cell[3]) {
    const double eps = 10E-6;
    bool inside = true;
    for (int a = 0; a < 3; a++) inside = inside && o.start[a] <= cell[a] && cell[a] <= o.end[a];
    if (inside) return true;
    double sl[3], dir[3];
    for (int a = 0; a < 3; a++) {
        sl[a] = static_cast<double>(src[a]) + 0.5;
        dir[a] = (static_cast<double>(cell[a]) + 0.5) - sl[a];
    }
    for (int axis = 0; axis < 3; axis++) {
        double sa[3], ea[3];
        for (int a = 0; a < 3; a++) {
            sa[a] = static_cast<double>(o.start[a]) + 0.5;
            ea[a] = static_cast<double>(o.end[a]) + 0.5;
        }
        for (int side = 0; side < 2; side++) {
            if (side == 1) std::swap(sa, ea);
            double nv[3] = {0, 0, 0};
            nv[axis] = 1;
            const double dot = nv[0] * dir[0] + nv[1] * dir[1] + nv[2] * dir[2];
            if (std::fabs(dot) > eps) {
                const double lambda = (sa[axis] - sl[axis]) / dir[axis];
                if (lambda <= 1 && lambda >= 0) {
                    double p[3];
                    for (int a = 0; a < 3; a++) p[a] = dir[a] * lambda + sl[a];
                    for (int a = 0; a < 3; a++) p[a] -= 0.5;
                    if (is_obstacle_cell_real(o, p)) return true;
                }
            }
        }
    }
    return false;
}
}  // namespace

std::vector<double> Domain::gauss_spatial(const GaussSettings &g) const {
    const LevelDesc &L = m_level[0];
    std::vector<double> out(static_cast<size_t>(L.size), 0.0);
    const double lo[3] = {m_settings.X[0], m_settings.Y[0], m_settings.Z[0]};
    double rs[3];
    size_t src[3];
    for (int a = 0; a < 3; a++) {
        rs[a] = g.sigma[a];
        rs[a] *= g.sigma[a];
        rs[a] *= 2;
        rs[a] = 1. / rs[a];
        src[a] = static_cast<size_t>(std::round((-lo[a] + g.pos[a]) / L.h[a]) + 1);  // Utility::get_index
    }
    const int n_obst = obstacle_count();
    const std::vector<uint64_t> inner = inner_list(0);
    double V = 0.;
    for (int pass = 0; pass < 2; pass++) {
        const double HRRrV = pass ? g.hrr / V : 0.0;
        const double rcp = pass ? 1. / g.cp : 0.0;
        for (uint64_t idx : inner) {
            size_t c[3];
            c[2] = idx / (static_cast<size_t>(L.N[0]) * L.N[1]);
            c[1] = (idx - c[2] * L.N[0] * L.N[1]) / L.N[0];
            c[0] = idx - c[2] * L.N[0] * L.N[1] - c[1] * L.N[0];
            bool blocked = false;
            for (int id = 0; id < n_obst && !blocked; id++) blocked = blocks_sight(m_obst[0][id], src, c);
            if (blocked) continue;
            double d[3];
            for (int a = 0; a < 3; a++) {
                const double mid = lo[a] + (static_cast<double>(c[a]) - 0.5) * L.h[a];
                d[a] = mid - g.pos[a];
            }
            for (int a = 0; a < 3; a++) d[a] *= d[a];
            for (int a = 0; a < 3; a++) d[a] *= rs[a];
            const double expr = std::exp(-(d[0] + d[1] + d[2]));
            if (pass == 0) V += expr * L.h[0] * L.h[1] * L.h[2];
            else out[idx] = HRRrV * rcp * expr;
        }
    }
    return out;
}

// source/Cube.cpp:35-80.  The reference computes the last index of the box
// from the domain's end coordinate (Utility::get_index(x_end, dx, X2)),
// which gives an index at or below 1 for every box inside the domain and a
// negative value converted to size_t otherwise, so its loop sets no cell or
// runs past the array.  Here the box covers the cells whose centres lie in
// it, by the rule the obstacles use (Obstacle.cpp: round((x1 - X1) / dx) + 1
// .. round((x2 - X1) / dx)), clipped to the inner cells; cells inside
// obstacles keep zero.
std::vector<double> Domain::cube_source(const CubeSettings &c) const {
    const LevelDesc &L = m_level[0];
    std::vector<double> out(static_cast<size_t>(L.size), 0.0);
    const double lo[3] = {m_settings.X[0], m_settings.Y[0], m_settings.Z[0]};
    int s[3], e[3];
    for (int a = 0; a < 3; a++) {
        s[a] = std::max(1, static_cast<int>(std::round((c.start[a] - lo[a]) / L.h[a])) + 1);
        e[a] = std::min(L.n[a], static_cast<int>(std::round((c.end[a] - lo[a]) / L.h[a])));
    }
    for (int k = s[2]; k <= e[2]; k++)
        for (int j = s[1]; j <= e[1]; j++)
            for (int i = s[0]; i <= e[0]; i++)
                if (!in_obstacle(L, i, j, k))
                    out[static_cast<size_t>(i) + static_cast<size_t>(L.N[0]) * (j + static_cast<size_t>(L.N[1]) * k)] =
                        c.value;
    return out;
}

void Domain::timer_clear() {
    for (auto &e : m_timer_events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    m_timer_events.clear();
}

void Domain::timer_arm(const std::string &tag) {
    require_device();
    ARTSS_HIP(hipStreamSynchronize(m_stream));
    timer_clear();
    m_timer_count = 0;
    m_timer_ms = 0.0;
    m_timer_tag = tag;
}

void Domain::timer_add(const EventPairs &pairs) {
    for (const auto &e : pairs) {
        float x = 0.f;
        ARTSS_HIP(hipEventElapsedTime(&x, e.first, e.second));
        m_timer_ms += x;
        m_timer_count++;
    }
}

void Domain::timer_read(long long *launches, double *total_ms) {
    require_device();
    ARTSS_HIP(hipStreamSynchronize(m_stream));
    timer_add(m_timer_events);
    *launches = m_timer_count;
    *total_ms = m_timer_ms;
    timer_clear();
    m_timer_count = 0;
    m_timer_ms = 0.0;
    m_timer_tag.clear();
}

double *Domain::alloc_field(int level) {
    require_device();
    double *p = nullptr;
    const size_t bytes = sizeof(double) * static_cast<size_t>(m_level[level].size);
    ARTSS_HIP(hipMalloc(&p, bytes));
    ARTSS_HIP(hipMemsetAsync(p, 0, bytes, m_stream));
    return p;
}

void Domain::free_field(double *p) {
    if (p && m_stream) {
        (void)hipStreamSynchronize(m_stream);
        (void)hipFree(p);
    }
}

}  // namespace artss
