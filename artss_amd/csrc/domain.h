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
    void apply_boundary(double *f, int field_type, int level);

    // fixed-order sum of `n` partials, blocking (one 8-byte readback)
    double reduce(const double *partial, int n);
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

    // Live kernel timing: while a tag is armed, every launch issued through
    // timed(tag, level, ...) is bracketed by HIP events on the domain stream.
    void timer_arm(const std::string &tag);
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
    long long m_face_entries = 0;
    FaceEntry *m_faces_dev = nullptr;
    FaceParam *m_params_dev = nullptr;

    std::string m_timer_tag;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> m_timer_events;
    void timer_clear();

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
