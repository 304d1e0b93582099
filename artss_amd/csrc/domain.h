// Geometry, boundary tables and device residency for one simulation domain.
//
// Host-side equivalent of the reference's DomainData (domain/DomainData.cpp),
// DomainController + Multigrid (domain/Multigrid.cpp) and the GPULists
// (GPULists/*JoinedList.cpp).  Instead of uploading sorted 8-byte index
// lists of every cell, a level is described by its box and obstacle boxes
// (device.h); only obstacle faces, which carry the reference's
// circular-constraint edits, go to HBM as explicit lists, already scheduled
// into conflict-free launch waves.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.h"
#include "settings.h"

namespace artss {

class hip_error : public std::runtime_error {
  public:
    explicit hip_error(const std::string &m) : std::runtime_error(m) {}
};

void hip_check(hipError_t e, const char *what);
#define ARTSS_HIP(call) ::artss::hip_check((call), #call)

struct ObstacleLevel {
    size_t start[3], end[3];
    std::vector<uint64_t> face[6];  // Obstacle::m_boundary per patch, sorted
};

// Step capture (Solver::set_step_graph).  While `mode` is kRecord, every
// residual-target loop of a step appends the iterations it took to `counts`
// (a loop that cannot be captured, or that ended other than by meeting its
// test, sets `failed`); while it is kCapture, loop number `next` enqueues
// budgets[next] iterations behind its device gate, writes its outcome to
// log[next] and reads nothing back to the host.
struct StepCapture {
    enum { kOff = 0, kRecord = 1, kCapture = 2 };
    static constexpr int kMaxLoops = 16;
    int mode = kOff;
    int next = 0;
    bool failed = false;
    std::vector<int> counts, budgets;
    std::vector<int> kinds;  // per loop: kDiffusion or kPressure
    enum { kDiffusion = 0, kPressure = 1 };
    LoopLog *log = nullptr;  // device, kMaxLoops entries, owned by the solver
};

// thrown by a loop asked to capture a form it has no capturable variant of
class capture_unsupported : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

class Domain {
  public:
    // device < 0: host-only (geometry and index lists; no stream, no device memory)
    Domain(const Settings &s, int device);
    ~Domain();
    Domain(const Domain &) = delete;
    Domain &operator=(const Domain &) = delete;

    const Settings &settings() const { return m_settings; }
    int levels() const { return m_levels; }
    const LevelDesc &level(int l) const { return m_level[l]; }
    hipStream_t stream() const { return m_stream; }
    int device() const { return m_device; }

    // DomainController::apply_boundary (DomainController.cpp:145): obstacle
    // faces first, then the six domain patches.
    // gate (optional): the fill's kernels do nothing while *gate != 0 (k::IterGate)
    void apply_boundary(double *f, int field_type, int level, const int *gate = nullptr);

    // A fill is simple when no write of it reads a cell that any of its
    // writes changes: it then needs one wave of face writes, faces and
    // ghosts can share one launch, and applying it twice equals applying it
    // once.
    bool bc_simple(int level, int field_type) const { return m_simple[level][field_type] != 0; }
    // An idempotent fill: no face write reads a cell the fill writes (it
    // needs one wave of faces, its result depends only on cells it does not
    // write, and applying it twice equals applying it once).  Kernels can
    // then produce the filled values themselves from bc_view.
    bool bc_idempotent(int level, int field_type) const { return m_idem[level][field_type] != 0; }
    // homogeneous: the same fill with every boundary value zero (its linear part)
    BCView bc_view(int level, int field_type, bool homogeneous = false) const;
    // the fill has no non-zero boundary value (it is linear)
    bool bc_values_zero(int level, int field_type) const;
    long long ghost_cells(int level) const;
    // apply_boundary as one launch; only for simple fills
    void apply_boundary_merged(double *f, int field_type, int level);
    // apply_boundary with every boundary value zero (the fill's linear part)
    void apply_boundary_linear(double *f, int field_type, int level);
    // the obstacle-face part of apply_boundary alone
    void apply_faces(double *f, int field_type, int level);

    // boundary tables, for kernels that apply the fill themselves
    const DomainBC &domain_bc(int level, int field_type) const { return m_dom_bc[level][field_type]; }
    int face_waves(int level, int field_type) const { return static_cast<int>(m_waves[level][field_type].size()); }
    long long face_wave_offset(int level, int field_type, int w) const {
        return m_waves[level][field_type][w].offset;
    }
    long long face_wave_count(int level, int field_type, int w) const { return m_waves[level][field_type][w].count; }
    const FaceEntry *faces_dev() const { return m_faces_dev; }
    const FaceParam *params_dev(int level, int field_type) const {
        return m_params_dev ? m_params_dev + m_param_offset[level][field_type] : nullptr;
    }

    // fixed-order sum of `n` partials, blocking (one 8-byte readback)
    double reduce(const double *partial, int n);
    // the same, enqueued only: the sum lands in pinned host memory, readable
    // with reduced_value() once the stream has been synchronised
    void reduce_async(const double *partial, int n);
    double reduced_value() const { return *m_sum_host; }
    double *partial_buffer() const { return m_partial; }
    int partial_capacity() const { return m_partial_cap; }

    // reference index lists, rebuilt on the host for parity checks
    std::vector<uint64_t> inner_list(int level) const;
    std::vector<uint64_t> domain_list(int level) const;
    std::vector<uint64_t> obstacle_list(int level) const;
    const ObstacleLevel &obstacle(int level, int id) const { return m_obst[level][id]; }
    long long face_entries() const { return m_face_entries; }
    int obstacle_count() const { return static_cast<int>(m_settings.obstacles.size()); }
    size_t inner_cells(int level) const;

    // GaussFunction::create_spatial_values (GaussFunction.cpp:52-109), host
    std::vector<double> gauss_spatial(const GaussSettings &g) const;
    // source/Cube.cpp:35-80: `value` on the inner cells of the box; see domain.cpp
    std::vector<double> cube_source(const CubeSettings &c) const;

    // Live kernel timing: while a tag is armed, every launch issued through
    // timed(tag, level, ...) is bracketed by HIP events on the domain stream.
    // (Work replayed from HIP graphs is not bracketed: owners of graphs run
    // eagerly while a tag is armed.)
    using EventPairs = std::vector<std::pair<hipEvent_t, hipEvent_t>>;
    void timer_arm(const std::string &tag);
    const std::string &timer_tag() const { return m_timer_tag; }
    StepCapture &capture() { return m_capture; }
    template <typename F>
    void timed(const char *kernel, int level, F &&launch) {
        if (m_timer_tag.empty() || m_timer_tag != std::string(kernel) + "@" + std::to_string(level)) {
            launch();
            return;
        }
        hipEvent_t a, b;
        ARTSS_HIP(hipEventCreate(&a));
        ARTSS_HIP(hipEventCreate(&b));
        ARTSS_HIP(hipEventRecord(a, m_stream));
        launch();
        ARTSS_HIP(hipEventRecord(b, m_stream));
        m_timer_events.emplace_back(a, b);
    }
    // launches recorded since timer_arm and their summed duration (ms); blocks
    void timer_read(long long *launches, double *total_ms);

    // device allocation helpers (zero-initialised)
    double *alloc_field(int level);
    void free_field(double *p);

  private:
    Settings m_settings;
    int m_device = 0;
    int m_levels = 0;
    std::vector<LevelDesc> m_level;
    std::vector<size_t> m_inner_count;  // fluid cells per level
    std::vector<std::vector<ObstacleLevel>> m_obst;  // [level][id]
    hipStream_t m_stream = nullptr;

    // domain ghost fill tables [level][field type]
    std::vector<std::vector<DomainBC>> m_dom_bc;
    // obstacle face waves [level][field type] -> list of (offset, count) into m_faces_dev
    struct Wave {
        long long offset, count;
    };
    std::vector<std::vector<std::vector<Wave>>> m_waves;
    std::vector<std::vector<long long>> m_param_offset;  // into m_params_dev
    std::vector<std::vector<char>> m_simple;             // [level][field type]
    std::vector<std::vector<char>> m_idem;               // [level][field type]
    std::vector<std::vector<char>> m_shell;              // [level][field type]: the fill writes every ghost
    std::vector<std::vector<char>> m_face_values_zero;   // [level][field type]
    FaceParam *m_params_hom_dev = nullptr;
    std::vector<std::vector<long long>> m_view_bits, m_view_deps, m_view_ndeps, m_view_rest, m_view_nrest;
    uint32_t *m_rest_dev = nullptr;
    uint32_t *m_bits_dev = nullptr;
    DepGhost *m_deps_dev = nullptr;
    long long m_face_entries = 0;
    FaceEntry *m_faces_dev = nullptr;
    FaceParam *m_params_dev = nullptr;

    std::string m_timer_tag;
    StepCapture m_capture;
    EventPairs m_timer_events;
    long long m_timer_count = 0;
    double m_timer_ms = 0.0;
    void timer_clear();
    void timer_add(const EventPairs &pairs);

    double *m_partial = nullptr;
    int m_partial_cap = 0;
    double *m_sum_dev = nullptr;
    double *m_sum_host = nullptr;

    void require_device() const;
    void build_levels();
    void build_obstacles();
    void build_boundary_tables();
};

}  // namespace artss
