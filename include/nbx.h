/*
 * nbx.h -- C ABI of the MI355X (gfx950) Barnes-Hut disk step.
 *
 * Drop-in for the reference's step-loop call surface (ualiberik/nbody-sim):
 * every entry point below names the reference function it replaces.  The
 * ABI uses plain pointers and sizes only; all particle pointers are DEVICE
 * pointers (HBM) unless a name says otherwise.
 *
 * Conventions
 *  - Return value: 0 on success, a negative NBX_E_* code on failure.  The
 *    message of the last failure on the calling thread is nbx_last_error()
 *    (the reference throws std::runtime_error from nbody_check_cuda,
 *    ParticleData.h:9-12; the C++ mirror in nbx.hpp rethrows it).
 *  - Every entry point returns after its work has completed on the device
 *    (the reference synchronises after each launch, BHTree.cuh:42), unless
 *    nbx_tree_set_async() switched the tree's stream to asynchronous mode.
 *  - Units: G = 1, M_sun = 1, AU, T0 = 1/(2 pi) yr (Perturber.h:13-18).
 *  - Not re-entrant on one nbx_tree: concurrent calls on the same tree race.
 */
#ifndef NBX_H
#define NBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBX_OK 0
#define NBX_E_ARG (-1)      /* invalid argument (null pointer, n mismatch, n <= 0)    */
#define NBX_E_HIP (-2)      /* HIP runtime error (launch, copy, allocation)           */
#define NBX_E_OVERFLOW (-3) /* a walk stack or node pool overflowed (no silent loss)  */
#define NBX_E_STATE (-4)    /* call order violated (e.g. force before build)          */

/* ParticleData (include/ParticleData.h:19-32): SoA, 10 float arrays + id.
 * Positions AU, velocities AU/T0, accelerations AU/T0^2, mass M_sun. */
typedef struct nbx_particles {
    float *x, *y, *z;
    float *vx, *vy, *vz;
    float *ax, *ay, *az;
    float *mass;
    int32_t *id;
    int32_t n;
} nbx_particles;

/* OctreeData (include/BHTree.cuh:13-36): opaque.  Owns the tree scratch
 * (keys, order, nodes, collision deltas), a HIP stream and, for the fused
 * step, a captured HIP graph. */
typedef struct nbx_tree nbx_tree;

/* Message of the last failed call on this thread ("" if none). */
const char *nbx_last_error(void);

/* ---- device memory for ParticleData (ParticleData.h:35-123) -------------- */
/* allocateParticlesGPU (ParticleData.h:61-75): all 11 arrays, zeroed. */
int nbx_particles_alloc(nbx_particles *p, int32_t n);
/* freeParticlesGPU (ParticleData.h:77-87). */
int nbx_particles_free(nbx_particles *p);
/* copyToGPU (ParticleData.h:90-105): host -> device, all 11 arrays. */
int nbx_copy_to_device(const nbx_particles *host, nbx_particles *dev);
/* copyToCPU (ParticleData.h:108-123): device -> host, all 11 arrays. */
int nbx_copy_to_host(const nbx_particles *dev, nbx_particles *host);

/* ---- tree lifetime (BHTree.cu:358-420) ----------------------------------- */
/* allocateOctree (BHTree.cu:358-381). */
int nbx_tree_create(nbx_tree **t, int32_t n_bodies);
/* Batched variant (no reference counterpart; the perturber-sweep batching):
 * n_seg independent disks stored back to back in one nbx_particles, disk s
 * holding seg_counts[s] bodies.  Every tree entry point then processes all
 * disks in one launch per phase; disk s never interacts with disk s'. */
int nbx_tree_create_batch(nbx_tree **t, int32_t n_seg, const int32_t *seg_counts);
/* freeOctree (BHTree.cu:383-395). */
int nbx_tree_destroy(nbx_tree *t);
/* resetOctree (BHTree.cu:397-420): clears the collision deltas; the Morton
 * build needs no pool reset. */
int nbx_tree_reset(nbx_tree *t);
/* Stream control: `stream` is a hipStream_t (NULL = the tree's own stream,
 * which the tree creates and keeps until nbx_tree_destroy; a caller's
 * stream is never destroyed by the tree).
 * async=1 makes entry points return without waiting for the device. */
int nbx_tree_set_stream(nbx_tree *t, void *stream);
/* No reference counterpart (accuracy option): multipole order of the tree
 * walk, 1 = monopole (the reference, BHTree.cu:581-583, default) or 2 = add
 * each accepted node's traceless quadrupole.  The opening rule and the
 * accepted set are unchanged.  Invalidates the current tree. */
int nbx_tree_set_multipole(nbx_tree *t, int order);
int nbx_tree_set_async(nbx_tree *t, int async);
int nbx_tree_sync(nbx_tree *t);

/* ---- tree build (BHTree.cuh:43-46) --------------------------------------- */
/* launchBBoxKernel (BHTree.cu:422-452): AABB, 1.001 padding and the root
 * cube of BHTree.cu:187-191, computed on the device. */
int nbx_bbox(nbx_tree *t, const nbx_particles *p);
/* launchBuildKernel (BHTree.cu:454-499): 63-bit keys from the root-cube
 * octant descent, stable radix sort, hierarchy from the sorted keys. */
int nbx_build(nbx_tree *t, const nbx_particles *p);
/* launchSummarizeKernel (BHTree.cu:501-516): node mass / COM, full
 * bottom-up reduction (no pass cap). */
int nbx_summarize(nbx_tree *t);
/* launchSortKernel (BHTree.cu:518-525): the leaf order.  The Morton sort of
 * nbx_build already is the DFS order; this call only validates the tree. */
int nbx_sort(nbx_tree *t);
/* OctreeData::sort_idx (BHTree.cuh:20): copy the body order to dev_out[n]. */
int nbx_sort_idx(nbx_tree *t, int32_t *dev_out);
/* Morton keys of the bodies in sorted order (dev_out[n], uint64). */
int nbx_sorted_keys(nbx_tree *t, uint64_t *dev_out);

/* ---- forces and collisions (BHTree.cuh:47-50, Perturber.h:73-75) --------- */
/* launchForceKernel (BHTree.cu:600-609): a = star term + tree self-gravity,
 * written with '='.  theta is the half-width / COM-distance ratio of
 * BHTree.cu:577.  self_a (optional, device [3][n] or NULL) receives the
 * tree part alone. */
int nbx_force(nbx_tree *t, nbx_particles *p, float theta, float softening, float star_mass);
int nbx_force_ex(nbx_tree *t, nbx_particles *p, float theta, float softening, float star_mass,
                 float *self_ax, float *self_ay, float *self_az);
/* No reference counterpart (diagnostic): per-body counts of accepted records
 * (self excluded) and opened internal records of the theta walk on the last
 * summarized tree, written to device arrays in original body order. */
int nbx_walk_counts(nbx_tree *t, float theta, int32_t *dev_n_acc, int32_t *dev_n_open);
/* No reference counterpart (diagnostic): statistics of the theta walk with the
 * tree's current shape, host out[4] = {mean accepted per body, mean opened
 * per body, mean records evaluated per wavefront, deepest wave stack}. */
int nbx_walk_stats(nbx_tree *t, float theta, double *out);
/* No reference counterpart (diagnostic): per-wavefront work of the same walk,
 * host out[3 * wave] = {records evaluated, record batches, deepest stack}.
 * *n_waves receives the wavefront count; out may be NULL to query it, else
 * it must hold cap >= *n_waves wavefronts. */
int nbx_walk_wave_stats(nbx_tree *t, float theta, int32_t *out, int32_t cap, int32_t *n_waves);
/* No reference counterpart (diagnostic): out[0] = colliding pairs counted by
 * the collision passes, out[1] = per-lane searches (collision / FoF) that
 * went on below a record without the per-lane stack (a full stack), both
 * since the tree's creation or the last clear; clear=1 zeroes them. */
int nbx_search_stats(nbx_tree *t, int32_t *out, int clear);
/* launchCollisionKernel (BHTree.cu:688-708): exact pair set
 * dist^2 < r_coll^2, impulse J = -(1+e) dvn / 2 for approaching pairs,
 * deltas accumulated per body in a fixed order, then v += dv. */
int nbx_collide(nbx_tree *t, nbx_particles *p, float r_coll, float restitution);
/* launchPerturberForceKernel (Perturber.cu:148-160): a += softened pull. */
int nbx_perturber_force(nbx_particles *p, float px, float py, float pz, float pmass,
                        float softening);

/* ---- integrator (Integrator.cuh:9-22) ------------------------------------ */
int nbx_half_kick(nbx_particles *p, float dt_half);          /* Integrator.cu:80-85  */
int nbx_drift(nbx_particles *p, float dt);                   /* Integrator.cu:87-92  */
int nbx_reset_acceleration(nbx_particles *p);                /* Integrator.cu:94-98  */
int nbx_gas_drag(nbx_particles *p, float drag_rate, float dt,
                 float star_mass);                           /* Integrator.cu:100-109 */

/* ---- Friends-of-Friends (Aggregate.h:17-18) ------------------------------ */
/* Partition of detectAggregates (Aggregate.cpp:26-76) for link length ll on
 * the device: dev_labels[i] = smallest particle index in i's group (1-body
 * groups label themselves).  Builds its own tree from p. */
int nbx_fof(nbx_tree *t, const nbx_particles *p, float link_length, int32_t *dev_labels);
/* No reference counterpart (batching): nbx_fof with link_lengths[s] for
 * disk s of a batched tree (one entry per disk; the reference links at each
 * run's own r_coll, main.cu:148).  Disks never link to each other. */
int nbx_fof_disks(nbx_tree *t, const nbx_particles *p, const float *link_lengths,
                  int32_t *dev_labels);

/* No counterpart in the reference: nbx_fof for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same labels as nbx_fof, with no tree: the scratch is a
 * static device array of 40 MiB per GPU (allocated when the library is
 * loaded, held until the process ends), so at most NBX_NOALLOC_MAX_BODIES
 * bodies (larger n: NBX_E_ARG before any device work).  No hipMalloc,
 * hipFree or other allocating call.  Work goes on `stream` (a hipStream_t;
 * NULL = the null stream) and the call returns after it completed.  Every
 * call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (other threads, other streams) race and must be serialized by the
 * caller. */
#define NBX_NOALLOC_MAX_BODIES 65536
int nbx_fof_noalloc(const nbx_particles *p, float link_length, int32_t *dev_labels, void *stream);

/* ---- fused step (main.cu:101-140) ---------------------------------------- */
typedef struct nbx_step_params {
    float dt, theta, softening, star_mass;
    float restitution, gas_drag_rate;
    int32_t collisions;      /* 1 = two-pass collision impulse each step          */
    int32_t use_graph;       /* 1 = replay one captured hipGraph per step         */
} nbx_step_params;

/* Per-disk values.  Every entry point taking `const nbx_disk_params *dp`
 * reads exactly one entry per disk of the tree: dp must point to n_seg
 * entries for a tree of nbx_tree_create_batch (n_seg disks), one for a
 * plain tree.  The length cannot be checked through a pointer; a shorter
 * array is read past its end. */
typedef struct nbx_disk_params {
    float r_coll;            /* collision_radius_factor * particlePhysicalRadius */
    int32_t perturber_enabled;
    float pert_mass_msun, pert_radius_au, pert_phase0, pert_omega;
} nbx_disk_params;

/* Initial forces of main.cu:85-96 (tree, star, perturber at t=0). */
int nbx_init_forces(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                    const nbx_disk_params *dp);
/* n_steps KDK steps of main.cu:101-140 starting at step index step0
 * (sim_time after step k is float(k+1)*dt).  Perturber positions follow
 * updatePerturber (Perturber.h:60-66) evaluated on the host. */
int nbx_step(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
             const nbx_disk_params *dp, int32_t step0, int32_t n_steps);
/* No reference counterpart (batching): n_steps steps of n_trees trees (each
 * a disk or a batch of disks, with its own particles ps[i] and disk
 * parameters dps[i]) stepped together: every tree replays its own captured
 * step on its own stream, step by step, so one tree's build passes run
 * beside another tree's force walk.  Each tree ends in exactly the state
 * nbx_step gives it alone.  The group starts after the work already on
 * trees[0]'s stream, and that stream waits for the whole group; without
 * use_graph, or with phase timing on any tree, the trees are stepped one
 * after the other.  A tree may appear once, and no two trees may step the
 * same particle arrays (NBX_E_ARG: they would race). */
#define NBX_MAX_GROUP 8
int nbx_step_concurrent(nbx_tree *const *trees, nbx_particles *const *ps, int32_t n_trees,
                        const nbx_step_params *sp, const nbx_disk_params *const *dps,
                        int32_t step0, int32_t n_steps);

/* ---- walk shape and autotuner (no reference counterpart) ----------------- */
/* Key sort algorithm: 0 = splitter buckets (sampled splitters, per-bucket
 * sort in LDS; the default for disks of at most 2^20 bodies), or the digit
 * width of the radix passes: 8, 9 or 11 bits (8, 7 or 6 passes over the
 * 63-bit keys; the default above 2^20 bodies).  The sorted order (ascending
 * key, ties by body index) does not depend on it. */
int nbx_tree_set_radix_bits(nbx_tree *t, int bits);
/* Targets per wavefront of the gravity walk (64, 32, 16 or 8; the 64 lanes
 * of a wave split each batch of tree records over 64/targets slots).  Results
 * depend on the shape only through the summation order. */
int nbx_tree_set_walk_shape(nbx_tree *t, int targets_per_wave);
/* Launch order of the gravity walk's wavefronts (results do not depend on
 * it): NBX_WAVES_LAUNCH = Morton order; NBX_WAVES_LONGEST_FIRST = the
 * stepping walk starts the waves that cost most in the previous step first;
 * NBX_WAVES_XCD (default) = each of the 8 accelerator dies (XCDs) takes one
 * contiguous Morton range of waves, so the tree records a die's waves read
 * stay in that die's L2. */
#define NBX_WAVES_LAUNCH 0
#define NBX_WAVES_LONGEST_FIRST 1
#define NBX_WAVES_XCD 2
int nbx_tree_set_wave_order(nbx_tree *t, int order);
int nbx_tree_get_walk_shape(nbx_tree *t, int32_t *targets_per_wave);
/* Times the gravity walk of every shape on the last summarized tree (outputs
 * go to scratch, the state is not touched), keeps the fastest, and writes
 * the mean ms per walk of shapes 64, 32, 16, 8 to ms_out[4] (may be NULL). */
int nbx_tune_walk(nbx_tree *t, float theta, float softening, float star_mass, float *ms_out);

/* ---- precision (no reference counterpart) -------------------------------- */
/* NBX_PRECISION_FLOAT (default): the reference's float step.
 * NBX_PRECISION_COMPENSATED: the KDK position / velocity updates (and the
 * drag) of nbx_step are evaluated in FP64 on hi + lo, the caller's float
 * arrays hold the rounded values and the tree holds the low-order parts;
 * forces stay in FP32.  Enabling clears the low parts (the state is its
 * float values).
 * NBX_PRECISION_FP64: the FP64 step, the reference of the precision study:
 * the tree holds the state in FP64; KDK, drag, node moments, the walk (same
 * octree and acceptance rule), star and perturber terms are FP64 operations;
 * the caller's float arrays receive the rounded state (positions,
 * velocities, accelerations).  nbx_init_forces loads the FP64 state from the
 * float arrays; nbx_step then advances it.  No collision pass, monopoles
 * only (NBX_E_ARG otherwise); at most 2^27 walk records. */
#define NBX_PRECISION_FLOAT 0
#define NBX_PRECISION_COMPENSATED 1
#define NBX_PRECISION_FP64 2
int nbx_tree_set_precision(nbx_tree *t, int mode);
/* No reference counterpart: the state as doubles into the device array
 * dev_out[9 n] = [x | y | z | vx | vy | vz | ax | ay | az] (body order): the
 * FP64 state, hi + lo in compensated mode (accelerations float), else the
 * float state. */
int nbx_state64(nbx_tree *t, const nbx_particles *p, double *dev_out);
/* FP64 sums over the state (low parts included in compensated mode, the
 * FP64 state in FP64 mode):
 * out[0] kinetic energy, out[1] softened star potential, out[2..4] angular
 * momentum, out[5] softened pairwise self potential (O(n^2), only if
 * with_self, else 0), out[6] total mass.  out is a host array of 7. */
int nbx_diagnostics(nbx_tree *t, const nbx_particles *p, float star_mass, float softening,
                    int with_self, double *out);

/* ---- reverse mode (no reference counterpart) ------------------------------
 * Adjoint of the collision-free step of main.cu:101-140 (KDK, gas drag,
 * monopole tree walk with the reference's acceptance, star, perturber), for
 * one disk per tree, float KDK.  The interaction lists are those of the
 * forward walk on the same positions.  Adjoint arrays are device arrays of
 * 3 n doubles laid out [x | y | z] in original body order; pgrad is a host
 * array of 2 doubles that accumulates dL/d(perturber mass in M_sun,
 * perturber radius in AU).  Calls return after the device work. */
/* Allocate the reverse-mode buffers; later builds also record the sorted
 * body range of every walk record. */
int nbx_tree_set_adjoint(nbx_tree *t, int enable);
/* xbar += (dA/dx)^T abar for A = tree + star + perturber acceleration at the
 * positions of p and the perturber at sim_time (rebuilds the tree). */
int nbx_adj_force(nbx_tree *t, const nbx_particles *p, float theta, float softening,
                  float star_mass, const nbx_disk_params *dp, float sim_time, const double *abar,
                  double *xbar, double *pgrad);
/* Reverse of step `step` (0-based, perturber at (step+1) dt): s0 is the
 * state (x, v, a) before the step, s1 after it.  On entry xbar, vbar are the
 * adjoints of x, v of s1 and abar the share of a of s1 from the next step's
 * first kick; on return they are the adjoints of x, v, a of s0. */
int nbx_adj_step(nbx_tree *t, const nbx_particles *s0, const nbx_particles *s1,
                 const nbx_step_params *sp, const nbx_disk_params *dp, int32_t step,
                 double *xbar, double *vbar, double *abar, double *pgrad);
/* Reverse of nbx_init_forces (perturber at t = 0): xbar += (dA/dx)^T abar at
 * the positions of s0. */
int nbx_adj_init_forces(nbx_tree *t, const nbx_particles *s0, const nbx_step_params *sp,
                        const nbx_disk_params *dp, const double *abar, double *xbar,
                        double *pgrad);

/* ---- run support on the host: config, disk, outputs ---------------------- */
/* Config (Config.h:5-35): the reference's 19 fields with its defaults, plus
 * generator fields with no reference counterpart: disk_seed (initDisk's
 * seed argument, default 42), and e_sigma / i_sigma (< 0, the default: the
 * reference's 1 % Gaussian velocity noise, DiskInit.cpp:55-60; >= 0:
 * Rayleigh eccentricities / inclinations of that width instead). */
typedef struct nbx_config {
    int32_t n_particles;
    float disk_r_min, disk_r_max, disk_h_factor;
    float particle_mass_msun, star_mass_msun, softening_AU, theta;
    float collision_radius_factor, restitution, gas_drag_rate;
    float dt, total_time;
    int32_t output_every_n_steps;
    char output_dir[512];
    int32_t perturber_enabled;
    float perturber_mass_mjup, perturber_radius_AU, perturber_phase0;
    uint32_t disk_seed;
    float e_sigma, i_sigma;
} nbx_config;
/* Defaults of Config.h:5-35. */
int nbx_config_default(nbx_config *cfg);
/* loadConfig (Config.cpp:5-55): JSON object, missing keys keep their
 * defaults, numbers parsed as double then narrowed, the checks of
 * Config.cpp:40-52 (NBX_E_ARG with the reference's message). */
int nbx_config_load(const char *path, nbx_config *cfg);
/* particlePhysicalRadius (DiskInit.cpp:10-16) of cfg->particle_mass_msun. */
float nbx_particle_radius(const nbx_config *cfg);
/* initDisk (DiskInit.cpp:18-67) into HOST arrays of cfg->n_particles bodies
 * (the same distribution; the random stream is this library's counter-based
 * generator keyed by seed, not std::mt19937). */
int nbx_init_disk(const nbx_config *cfg, uint32_t seed, nbx_particles *host);

/* Aggregate row (Aggregate.h:4-15 without the member list). */
typedef struct nbx_aggregate {
    int32_t id, n_particles;
    float mass_msun, cx, cy, cz, vx, vy, vz, dist_center;
} nbx_aggregate;
/* OutputWriter (OutputWriter.h / OutputWriter.cpp): frames.bin version 2
 * and stats.csv in `dir`.  Host arrays throughout. */
typedef struct nbx_writer nbx_writer;
/* OutputWriter::OutputWriter: create both files, write both headers. */
int nbx_writer_open(nbx_writer **w, const char *dir, int32_t n_particles, float dt,
                    int32_t output_every_n_steps);
/* writeFrame: time, x[n], y[n], z[n], agg_size[n] (uint16). */
int nbx_writer_frame(nbx_writer *w, const nbx_particles *host, float sim_time_T0,
                     const uint16_t *agg_sizes);
/* writeStats: one row per aggregate. */
int nbx_writer_stats(nbx_writer *w, int32_t frame, float sim_time_T0, int32_t n_aggregates,
                     const nbx_aggregate *aggs);
/* writePerturberStats: the agg_id = -1 row. */
int nbx_writer_perturber(nbx_writer *w, int32_t frame, float sim_time_T0, int32_t n_aggregates,
                         float px, float py, float pz, float mass_msun, float omega);
/* finalize: patch n_frames in the header and flush. */
int nbx_writer_finalize(nbx_writer *w, int32_t frames_written);
int nbx_writer_close(nbx_writer *w);
/* No reference counterpart in the engine: the visualizer's perturber.json
 * sidecar (PerturberRenderer.cs:67-73 field names). */
int nbx_write_perturber_sidecar(const char *dir, int32_t enabled, float mass_msun, float r_AU,
                                float phase0_rad, float omega);

/* ---- diagnostics --------------------------------------------------------- */
/* Tree statistics of the last build: number of octree nodes and maximum
 * depth per disk (host arrays of length n_seg, may be NULL). */
int nbx_tree_stats(nbx_tree *t, int32_t *n_nodes, int32_t *max_level);
/* Average duration (ms) of the named phase's kernels over the last
 * nbx_step call when timing was enabled with nbx_tree_set_timing(t, 1).
 * Phases: "kdk1", "build", "walk", "collide", "kdk2".  Returns 0 and
 * writes *ms, or NBX_E_ARG for an unknown phase. */
int nbx_tree_set_timing(nbx_tree *t, int on);
int nbx_phase_ms(nbx_tree *t, const char *phase, float *ms);
/* Achievable FP32 rate of this GPU (no reference counterpart; measurement
 * only): kind 0 times eight independent FMA chains per thread (2 flop per
 * FMA), kind 1 the walk's record evaluation on registers over an LDS table
 * (its distance, opening test, selects, rsqrt and accumulation, priced like
 * the walk at 20 flop per record).  Writes TFLOP/s. */
int nbx_probe_fp32(int kind, double *tflops);

#ifdef __cplusplus
}
#endif
#endif /* NBX_H */
