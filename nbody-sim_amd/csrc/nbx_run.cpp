// nbx_run.cpp -- the simulation driver of src/main.cu (config -> disk ->
// step loop -> frames.bin / stats.csv), written against the C++ mirror of
// the reference's call surface (include/nbx.hpp) over libnbx.so.
//
//   nbx_run [config.json] [--out DIR] [--steps N] [--plain] [--no-output]
//
// Without --out the run goes to <output_dir>/<YYYYmmdd_HHMMSS> as in
// main.cu:24-37; --steps caps the number of steps (for smoke runs).
// --plain steps with the reference's own call sequence (main.cu:85-138: one
// synchronous library call per kernel, as every launch function of the
// reference synchronizes) instead of the captured step; --no-output skips
// frames, statistics and FoF and prints the wall time of the step loop as
// the last line, {"n": N, "steps": K, "loop_s": T}.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nbx.hpp"

using namespace nbx_ref;

int main(int argc, char **argv)
{
    std::string cfg_path = "config/simulation.json", out_dir;
    long max_steps = -1;
    bool plain = false, output = true;
    for (int a = 1; a < argc; ++a) {
        if (!strcmp(argv[a], "--out") && a + 1 < argc) out_dir = argv[++a];
        else if (!strcmp(argv[a], "--steps") && a + 1 < argc) max_steps = atol(argv[++a]);
        else if (!strcmp(argv[a], "--plain")) plain = true;
        else if (!strcmp(argv[a], "--no-output")) output = false;
        else cfg_path = argv[a];
    }
    try {
        std::cout << "Loading config: " << cfg_path << "\n";
        Config cfg = loadConfig(cfg_path);
        if (out_dir.empty() && output) {  // timestamped run directory (main.cu:24-37)
            char ts[20];
            const std::time_t now = std::time(nullptr);
            std::tm tm_buf{};
            localtime_r(&now, &tm_buf);
            std::strftime(ts, sizeof(ts), "%Y%m%d_%H%M%S", &tm_buf);
            out_dir = std::string(cfg.output_dir) + "/" + ts;
        }
        if (output) {
            std::filesystem::create_directories(out_dir);
            std::cout << "Output dir: " << out_dir << "\n";
        }

        std::cout << "Initializing disk (" << cfg.n_particles << " particles)...\n";
        ParticleData cpu = initDisk(cfg, cfg.disk_seed);
        ParticleData gpu = allocateParticlesGPU(cfg.n_particles);
        copyToGPU(cpu, gpu);
        OctreeData tree = allocateOctree(cfg.n_particles);

        const float r_coll = cfg.collision_radius_factor * particlePhysicalRadius(cfg);
        Perturber pert = makePerturber(cfg);
        std::unique_ptr<OutputWriter> writer;
        if (output) {
            writer = std::make_unique<OutputWriter>(out_dir, cfg);
            if (pert.enabled)
                check(nbx_write_perturber_sidecar(out_dir.c_str(), 1, pert.mass_msun, pert.r,
                                                  pert.phase0, pert.omega),
                      "perturber sidecar");
        }

        long total_steps = static_cast<long>(cfg.total_time / cfg.dt);  // main.cu:64
        if (max_steps >= 0) total_steps = std::min(total_steps, max_steps);
        const int every = cfg.output_every_n_steps;

        nbx_step_params sp{cfg.dt, cfg.theta, cfg.softening_AU, cfg.star_mass_msun,
                           cfg.restitution, cfg.gas_drag_rate, 1, 1};
        nbx_disk_params dp{r_coll, pert.enabled ? 1 : 0, pert.mass_msun, pert.r, pert.phase0,
                           pert.omega};
        // the reference's tree rebuild and force evaluation (main.cu:85-91, :112-118)
        auto plain_forces = [&]() {
            resetOctree(tree);
            launchBBoxKernel(tree, gpu);
            launchBuildKernel(tree, gpu);
            launchSummarizeKernel(tree);
            launchSortKernel(tree);
            launchResetAccelerationKernel(gpu);
            launchForceKernel(tree, gpu, cfg.theta, cfg.softening_AU, cfg.star_mass_msun);
        };
        if (plain) {  // main.cu:85-96
            plain_forces();
            if (pert.enabled) {
                updatePerturber(pert, 0.0f);
                launchPerturberForceKernel(gpu, pert.x, pert.y, pert.z, pert.mass_msun,
                                           cfg.softening_AU);
            }
        } else {
            check(nbx_init_forces(tree.h, &gpu, &sp, &dp), "initial forces");
        }
        // one step of main.cu:104-138 through one library call per kernel
        auto plain_step = [&](long step) {
            launchHalfKickKernel(gpu, cfg.dt * 0.5f);
            launchDriftKernel(gpu, cfg.dt);
            const float sim_time_now = static_cast<float>(step + 1) * cfg.dt;
            plain_forces();
            if (pert.enabled) {
                updatePerturber(pert, sim_time_now);
                launchPerturberForceKernel(gpu, pert.x, pert.y, pert.z, pert.mass_msun,
                                           cfg.softening_AU);
            }
            launchCollisionKernel(tree, gpu, r_coll, cfg.restitution);
            launchHalfKickKernel(gpu, cfg.dt * 0.5f);
            launchGasDragKernel(gpu, cfg.gas_drag_rate, cfg.dt, cfg.star_mass_msun);
        };

        int frames_written = 0;
        const auto t_loop = std::chrono::steady_clock::now();
        for (long step0 = 0; step0 < total_steps;) {
            // run to the next output step ((step + 1) % every == 0, main.cu:143)
            const long next = output ? std::min(total_steps, (step0 / every + 1) * every)
                                     : total_steps;
            if (plain) {
                for (long k = step0; k < next; ++k) plain_step(k);
            } else {
                check(nbx_step(tree.h, &gpu, &sp, &dp, (int32_t)step0, (int32_t)(next - step0)),
                      "step");
            }
            step0 = next;
            if (!output || step0 % every != 0) continue;
            const float sim_time_now = static_cast<float>(step0) * cfg.dt;
            copyToCPU(gpu, cpu);
            // --plain: the reference's own call of main.cu:148 (host state);
            // otherwise the partition reuses the step's tree scratch
            auto aggs = plain ? detectAggregates(cpu, r_coll, cfg)
                              : detectAggregates(tree, gpu, cpu, r_coll, cfg.particle_mass_msun);
            std::vector<uint16_t> agg_sizes(cfg.n_particles, 1);
            for (const Aggregate &a : aggs) {
                const uint16_t nc = static_cast<uint16_t>(std::min(a.n_particles, 65535));
                for (int pid : a.particle_ids)
                    if (pid >= 0 && pid < cfg.n_particles) agg_sizes[pid] = nc;
            }
            writer->writeFrame(cpu, sim_time_now, agg_sizes);
            writer->writeStats(aggs, frames_written, sim_time_now);
            if (pert.enabled) {
                updatePerturber(pert, sim_time_now);
                writer->writePerturberStats(frames_written, sim_time_now,
                                           static_cast<int>(aggs.size()), pert.x, pert.y,
                                           pert.z, pert.mass_msun, pert.omega);
            }
            ++frames_written;
            const float pct = 100.f * step0 / total_steps;
            std::cout << std::fixed << std::setprecision(1) << "\r[" << pct
                      << "%] t=" << sim_time_now << " T0  aggregates=" << aggs.size()
                      << "    " << std::flush;
        }
        if (output) {
            writer->finalize(frames_written);
            std::cout << "\nDone. " << frames_written << " frames written to " << out_dir << "\n";
        } else {
            check(nbx_tree_sync(tree.h), "sync");
            const double loop_s =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
            std::cout << "{\"n\": " << cfg.n_particles << ", \"steps\": " << total_steps
                      << ", \"loop_s\": " << std::setprecision(9) << loop_s << "}\n";
        }
        freeOctree(tree);
        freeParticlesGPU(gpu);
        freeParticlesCPU(cpu);
    } catch (const std::exception &e) {
        std::cerr << "\nSimulation error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
