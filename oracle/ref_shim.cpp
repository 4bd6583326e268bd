// ref_shim.cpp -- C entry points over the reference's own host-side C++
// (Aggregate.cpp, DiskInit.cpp, OutputWriter.cpp), compiled from /root/reference by
// oracle/Makefile into oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY:
// used to check the oracle's FoF / disk-init restatements, never by the
// product path.
#include "Aggregate.h"
#include "Config.h"
#include "DiskInit.h"
#include "OutputWriter.h"
#include "ParticleData.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// detectAggregates (Aggregate.h:17-18) on caller arrays.  Writes
// labels[i] = smallest index in i's group (singletons: i itself), and per
// aggregate (ordered by smallest member index) n, mass, com[3], vel[3],
// dist.  Returns the number of aggregates, or -(needed) if cap is too small.
int ref_detect_aggregates(int n, const float* x, const float* y, const float* z,
                          const float* vx, const float* vy, const float* vz,
                          float link_length, float particle_mass, int32_t* labels,
                          int cap, int32_t* agg_n, float* agg_f /* [cap*8] */) {
    ParticleData p = allocateParticlesCPU(n);
    std::memcpy(p.x, x, n * sizeof(float));
    std::memcpy(p.y, y, n * sizeof(float));
    std::memcpy(p.z, z, n * sizeof(float));
    std::memcpy(p.vx, vx, n * sizeof(float));
    std::memcpy(p.vy, vy, n * sizeof(float));
    std::memcpy(p.vz, vz, n * sizeof(float));
    for (int i = 0; i < n; ++i) p.id[i] = i;
    Config cfg;
    cfg.particle_mass_msun = particle_mass;
    std::vector<Aggregate> aggs = detectAggregates(p, link_length, cfg);
    freeParticlesCPU(p);
    for (int i = 0; i < n; ++i) labels[i] = i;
    std::sort(aggs.begin(), aggs.end(), [](const Aggregate& a, const Aggregate& b) {
        return *std::min_element(a.particle_ids.begin(), a.particle_ids.end()) <
               *std::min_element(b.particle_ids.begin(), b.particle_ids.end());
    });
    int na = static_cast<int>(aggs.size());
    if (na > cap) return -na;
    for (int k = 0; k < na; ++k) {
        const Aggregate& a = aggs[k];
        int mn = *std::min_element(a.particle_ids.begin(), a.particle_ids.end());
        for (int pid : a.particle_ids) labels[pid] = mn;
        agg_n[k] = a.n_particles;
        float* f = agg_f + 8 * k;
        f[0] = a.mass_msun;
        f[1] = a.cx; f[2] = a.cy; f[3] = a.cz;
        f[4] = a.vx; f[5] = a.vy; f[6] = a.vz;
        f[7] = a.dist_center;
    }
    return na;
}

// particlePhysicalRadius (DiskInit.h:13-14)
float ref_particle_physical_radius(float particle_mass) {
    Config cfg;
    cfg.particle_mass_msun = particle_mass;
    return particlePhysicalRadius(cfg);
}

// OutputWriter (OutputWriter.h:14-49) over a scripted run: n_frames frames
// of positions xyz[f][3][n] at times[f] with agg_sizes[f][n], after each
// frame the same n_agg aggregate rows (ai[k] = {id, n_particles}, af[k] =
// {mass, cx, cy, cz, vx, vy, vz, dist}) and, if pert, the perturber row.
int ref_write_run(const char* dir, int n, float dt, int every, int n_frames,
                  const float* xyz, const float* times, const uint16_t* agg_sizes,
                  int n_agg, const int32_t* ai, const float* af,
                  int pert, float px, float py, float pz, float pmass, float omega) {
    Config cfg;
    cfg.n_particles = n;
    cfg.dt = dt;
    cfg.output_every_n_steps = every;
    std::vector<Aggregate> aggs(n_agg);
    for (int k = 0; k < n_agg; ++k) {
        Aggregate& a = aggs[k];
        a.id = ai[2 * k];
        a.n_particles = ai[2 * k + 1];
        const float* f = af + 8 * k;
        a.mass_msun = f[0];
        a.cx = f[1]; a.cy = f[2]; a.cz = f[3];
        a.vx = f[4]; a.vy = f[5]; a.vz = f[6];
        a.dist_center = f[7];
    }
    ParticleData p = allocateParticlesCPU(n);
    {
        OutputWriter w(dir, cfg);
        for (int fr = 0; fr < n_frames; ++fr) {
            const float* b = xyz + (size_t)fr * 3 * n;
            std::memcpy(p.x, b, n * sizeof(float));
            std::memcpy(p.y, b + n, n * sizeof(float));
            std::memcpy(p.z, b + 2 * n, n * sizeof(float));
            std::vector<uint16_t> sizes(agg_sizes + (size_t)fr * n, agg_sizes + (size_t)(fr + 1) * n);
            w.writeFrame(p, times[fr], sizes);
            w.writeStats(aggs, fr, times[fr]);
            if (pert) w.writePerturberStats(fr, times[fr], n_agg, px, py, pz, pmass, omega);
        }
        w.finalize(n_frames);
    }
    freeParticlesCPU(p);
    return 0;
}

}  // extern "C"
