// nbx.hpp -- C++ mirror of the reference's launcher surface over the C ABI.
//
// A translation unit of the reference's host code (src/main.cu, the tests)
// that includes this header instead of BHTree.cuh / Integrator.cuh /
// Perturber.h / Aggregate.h keeps its call sites unchanged: the same type
// and function names, the same argument meaning, std::runtime_error on
// failure (ParticleData.h:9-12), synchronous at return.  Only the device
// memory comes from nbx_particles_alloc (HIP) instead of cudaMalloc.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbx.h"

namespace nbx_ref {

inline void check(int rc, const char *op)
{
    if (rc != NBX_OK) throw std::runtime_error(std::string(op) + ": " + nbx_last_error());
}

// Config (Config.h:5-35).
using Config = nbx_config;

// ParticleData (ParticleData.h:19-32): device pointers when allocated by
// allocateParticlesGPU, host pointers when allocated by allocateParticlesCPU.
using ParticleData = nbx_particles;

inline ParticleData allocateParticlesCPU(int n)  // ParticleData.h:35-44
{
    ParticleData p{};
    p.n = n;
    float **f[10] = {&p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.ax, &p.ay, &p.az, &p.mass};
    for (auto *q : f) *q = new float[n]();
    p.id = new int32_t[n]();
    return p;
}

inline void freeParticlesCPU(ParticleData &p)  // ParticleData.h:46-57
{
    float *f[10] = {p.x, p.y, p.z, p.vx, p.vy, p.vz, p.ax, p.ay, p.az, p.mass};
    for (auto *q : f) delete[] q;
    delete[] p.id;
    p = ParticleData{};
}

inline ParticleData allocateParticlesGPU(int n)  // ParticleData.h:61-75
{
    ParticleData p{};
    check(nbx_particles_alloc(&p, n), "allocateParticlesGPU");
    return p;
}

inline void freeParticlesGPU(ParticleData &p) { check(nbx_particles_free(&p), "freeParticlesGPU"); }

inline void copyToGPU(const ParticleData &cpu, ParticleData &gpu)  // ParticleData.h:90-105
{
    check(nbx_copy_to_device(&cpu, &gpu), "copyToGPU");
}

inline void copyToCPU(const ParticleData &gpu, ParticleData &cpu)  // ParticleData.h:108-123
{
    check(nbx_copy_to_host(&gpu, &cpu), "copyToCPU");
}

// OctreeData (BHTree.cuh:13-36): the tree scratch lives behind the handle.
struct OctreeData {
    nbx_tree *h = nullptr;
    int n_bodies = 0;
};

inline OctreeData allocateOctree(int n_bodies)  // BHTree.cu:358-381
{
    OctreeData t;
    check(nbx_tree_create(&t.h, n_bodies), "allocateOctree");
    t.n_bodies = n_bodies;
    return t;
}
inline void freeOctree(OctreeData &t) { check(nbx_tree_destroy(t.h), "freeOctree"); t.h = nullptr; }
inline void resetOctree(OctreeData &t) { check(nbx_tree_reset(t.h), "resetOctree"); }

inline void launchBBoxKernel(OctreeData &t, const ParticleData &p)
{
    check(nbx_bbox(t.h, &p), "launchBBoxKernel");
}
inline void launchBuildKernel(OctreeData &t, const ParticleData &p)
{
    check(nbx_build(t.h, &p), "launchBuildKernel");
}
inline void launchSummarizeKernel(OctreeData &t) { check(nbx_summarize(t.h), "launchSummarizeKernel"); }
inline void launchSortKernel(OctreeData &t) { check(nbx_sort(t.h), "launchSortKernel"); }
inline void launchForceKernel(OctreeData &t, ParticleData &p, float theta, float softening,
                              float star_mass)
{
    check(nbx_force(t.h, &p, theta, softening, star_mass), "launchForceKernel");
}
inline void launchCollisionKernel(OctreeData &t, ParticleData &p, float r_coll, float restitution)
{
    check(nbx_collide(t.h, &p, r_coll, restitution), "launchCollisionKernel");
}

// Integrator.cuh:9-22
inline void launchHalfKickKernel(ParticleData &g, float dt_half)
{
    check(nbx_half_kick(&g, dt_half), "launchHalfKickKernel");
}
inline void launchDriftKernel(ParticleData &g, float dt) { check(nbx_drift(&g, dt), "launchDriftKernel"); }
inline void launchResetAccelerationKernel(ParticleData &g)
{
    check(nbx_reset_acceleration(&g), "launchResetAccelerationKernel");
}
inline void launchGasDragKernel(ParticleData &g, float drag_rate, float dt, float star_mass_msun)
{
    check(nbx_gas_drag(&g, drag_rate, dt, star_mass_msun), "launchGasDragKernel");
}

// Perturber.h:27-75 (host-side orbit in float, device pull)
constexpr float MJUP_TO_MSUN = 9.5479e-4f;
struct Perturber {
    bool enabled = false;
    float r = 0.f, mass_msun = 0.f, phase0 = 0.f, omega = 0.f;
    float x = 0.f, y = 0.f, z = 0.f;
};
inline Perturber makePerturber(bool enabled, float mass_mjup, float radius_au, float phase0,
                               float star_mass_msun)
{
    Perturber p;
    p.enabled = enabled;
    if (!enabled) return p;
    p.r = radius_au;
    p.mass_msun = mass_mjup * MJUP_TO_MSUN;
    p.phase0 = phase0;
    p.omega = std::sqrt(star_mass_msun / (p.r * p.r * p.r));
    p.x = p.r * std::cos(p.phase0);
    p.z = p.r * std::sin(p.phase0);
    return p;
}
inline void updatePerturber(Perturber &p, float sim_time)
{
    if (!p.enabled) return;
    float phi = p.phase0 + p.omega * sim_time;
    p.x = p.r * std::cos(phi);
    p.z = p.r * std::sin(phi);
    p.y = 0.f;
}
inline void launchPerturberForceKernel(ParticleData &g, float px, float py, float pz, float pmass,
                                       float softening)
{
    check(nbx_perturber_force(&g, px, py, pz, pmass, softening), "launchPerturberForceKernel");
}

// Aggregate.h:4-18, on device particles: the partition comes from the GPU,
// the per-group sums follow Aggregate.cpp:58-69 on the host.
struct Aggregate {
    int id = 0, n_particles = 0;
    float mass_msun = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
    float dist_center = 0.f;
    std::vector<int> particle_ids;
};

// Aggregates of a partition: groups of >= 2 bodies (lab[i] = smallest index
// of i's group, host array), each with the float sums of Aggregate.cpp:58-69
// over its members in ascending index, from the host state; aggregates are
// numbered in order of their smallest member.
inline std::vector<Aggregate> aggregates_of_labels(const ParticleData &host, const int32_t *lab,
                                                   float particle_mass_msun)
{
    const int n = host.n;
    std::vector<int> size(n, 0);
    for (int i = 0; i < n; ++i) ++size[lab[i]];
    std::vector<Aggregate> out;
    std::vector<int> slot(n, -1);
    for (int i = 0; i < n; ++i) {  // roots are the smallest members: ascending order
        const int r = lab[i];
        if (size[r] < 2) continue;
        if (slot[r] < 0) {
            slot[r] = (int)out.size();
            Aggregate a;
            a.id = slot[r];
            a.n_particles = size[r];
            a.mass_msun = a.n_particles * particle_mass_msun;
            a.particle_ids.reserve(size[r]);
            out.push_back(std::move(a));
        }
        Aggregate &a = out[slot[r]];
        a.particle_ids.push_back(i);
        a.cx += host.x[i]; a.cy += host.y[i]; a.cz += host.z[i];
        a.vx += host.vx[i]; a.vy += host.vy[i]; a.vz += host.vz[i];
    }
    for (Aggregate &a : out) {
        float inv = 1.0f / static_cast<float>(a.n_particles);
        a.cx *= inv; a.cy *= inv; a.cz *= inv;
        a.vx *= inv; a.vy *= inv; a.vz *= inv;
        a.dist_center = std::sqrt(a.cx * a.cx + a.cy * a.cy + a.cz * a.cz);
    }
    return out;
}

// `gpu` holds device pointers, `host` a host copy of the same state (the
// reference calls detectAggregates after copyToCPU, main.cu:144-148): the
// partition uses the caller's tree scratch.
inline std::vector<Aggregate> detectAggregates(OctreeData &t, const ParticleData &gpu,
                                               const ParticleData &host, float link_length,
                                               float particle_mass_msun)
{
    // device label buffer: the id array of a scratch ParticleData
    ParticleData lab{}, lab_host{};
    try {
        check(nbx_particles_alloc(&lab, gpu.n), "detectAggregates labels");
        check(nbx_fof(t.h, &gpu, link_length, lab.id), "detectAggregates");
        lab_host = allocateParticlesCPU(gpu.n);
        check(nbx_copy_to_host(&lab, &lab_host), "detectAggregates labels");
        std::vector<Aggregate> out = aggregates_of_labels(host, lab_host.id, particle_mass_msun);
        freeParticlesCPU(lab_host);
        check(nbx_particles_free(&lab), "detectAggregates labels");
        return out;
    } catch (...) {
        freeParticlesCPU(lab_host);
        nbx_particles_free(&lab);
        throw;
    }
}

// detectAggregates (Aggregate.h:17-18) with the reference's own signature:
// `p` is a HOST ParticleData, as at main.cu:148 after copyToCPU.  The
// positions go to the device, where a tree of their own gives the
// partition (the reference's O(N^2) union-find, Aggregate.cpp:26-76, has
// the same components); the sums use cfg.particle_mass_msun.
inline std::vector<Aggregate> detectAggregates(const ParticleData &p, float link_length,
                                               const Config &cfg)
{
    ParticleData dev{};
    nbx_tree *h = nullptr;
    try {
        check(nbx_particles_alloc(&dev, p.n), "detectAggregates particles");
        check(nbx_copy_to_device(&p, &dev), "detectAggregates particles");
        check(nbx_tree_create(&h, p.n), "detectAggregates tree");
        OctreeData t;
        t.h = h;
        t.n_bodies = p.n;
        std::vector<Aggregate> out = detectAggregates(t, dev, p, link_length,
                                                      cfg.particle_mass_msun);
        check(nbx_tree_destroy(h), "detectAggregates tree");
        check(nbx_particles_free(&dev), "detectAggregates particles");
        return out;
    } catch (...) {
        if (h) nbx_tree_destroy(h);
        nbx_particles_free(&dev);
        throw;
    }
}

// ---- run support (Config.h, DiskInit.h, OutputWriter.h) -------------------

inline Config loadConfig(const std::string &path)  // Config.cpp:5-55
{
    Config c;
    check(nbx_config_load(path.c_str(), &c), "loadConfig");
    return c;
}

inline float particlePhysicalRadius(const Config &cfg) { return nbx_particle_radius(&cfg); }

inline ParticleData initDisk(const Config &cfg, unsigned int seed = 42)  // DiskInit.h:9
{
    ParticleData p = allocateParticlesCPU(cfg.n_particles);
    try {
        check(nbx_init_disk(&cfg, seed, &p), "initDisk");
    } catch (...) {
        freeParticlesCPU(p);
        throw;
    }
    return p;
}

inline Perturber makePerturber(const Config &cfg)  // Perturber.h:39-57
{
    return makePerturber(cfg.perturber_enabled != 0, cfg.perturber_mass_mjup,
                         cfg.perturber_radius_AU, cfg.perturber_phase0, cfg.star_mass_msun);
}

constexpr float T0_TO_YEARS = 1.0f / (2.0f * 3.14159265f);  // OutputWriter.h:10-12
constexpr float MSUN_TO_MEARTH = 332946.0f;

class OutputWriter {  // OutputWriter.h:14-49
public:
    OutputWriter(const std::string &output_dir, const Config &cfg) : n_(cfg.n_particles)
    {
        check(nbx_writer_open(&w_, output_dir.c_str(), cfg.n_particles, cfg.dt,
                              cfg.output_every_n_steps),
              "OutputWriter");
    }
    ~OutputWriter() { nbx_writer_close(w_); }
    OutputWriter(const OutputWriter &) = delete;
    OutputWriter &operator=(const OutputWriter &) = delete;
    void writeFrame(const ParticleData &cpu, float sim_time_T0, const std::vector<uint16_t> &agg)
    {
        if ((int)agg.size() != n_)
            throw std::runtime_error(
                "OutputWriter::writeFrame: agg_sizes_per_particle size mismatch");
        check(nbx_writer_frame(w_, &cpu, sim_time_T0, agg.data()), "writeFrame");
    }
    void writeStats(const std::vector<Aggregate> &aggs, int frame_idx, float sim_time_T0)
    {
        std::vector<nbx_aggregate> rows;
        rows.reserve(aggs.size());
        for (const Aggregate &a : aggs)
            rows.push_back(nbx_aggregate{a.id, a.n_particles, a.mass_msun, a.cx, a.cy, a.cz,
                                         a.vx, a.vy, a.vz, a.dist_center});
        check(nbx_writer_stats(w_, frame_idx, sim_time_T0, (int)rows.size(), rows.data()),
              "writeStats");
    }
    void writePerturberStats(int frame_idx, float sim_time_T0, int n_aggregates, float px,
                             float py, float pz, float mass_msun, float omega)
    {
        check(nbx_writer_perturber(w_, frame_idx, sim_time_T0, n_aggregates, px, py, pz,
                                   mass_msun, omega),
              "writePerturberStats");
    }
    void finalize(int total_frames_written)
    {
        check(nbx_writer_finalize(w_, total_frames_written), "finalize");
    }

private:
    nbx_writer *w_ = nullptr;
    int n_ = 0;
};

}  // namespace nbx_ref
