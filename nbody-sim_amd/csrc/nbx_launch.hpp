// nbx_launch.hpp -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include "nbx_internal.hpp"

namespace nbx {

// Device scratch owned by one nbx_tree.  Sizes are for n bodies over all
// disks; see nbx_internal.hpp for the layout.
struct Bufs {
    int n = 0;
    int n_sweep_tiles = 0;       // radix-sort tiles (kSweepTile keys each)
    int n_scan_tiles = 0;        // prefix-scan tiles (kScanTile bodies each)
    uint64_t *keys[2] = {};      // double-buffered sort keys
    int *vals[2] = {};           // double-buffered body indices
    int *ghist = nullptr;        // [kPasses * 256] global digit histograms
    unsigned *status[2] = {};    // [n_sweep_tiles * 256] look-back words per pass
    int *tickets = nullptr;      // [32] tile tickets (per pass / scan) + counters
    double4 *prefix = nullptr;   // [n + 1] exclusive prefix of {m, m x, m y, m z}
    int *pflag = nullptr;        // [n_scan_tiles] scan look-back flags
    double4 *pagg = nullptr;     // [n_scan_tiles] tile aggregates
    double4 *pinc = nullptr;     // [n_scan_tiles] inclusive tile prefixes
    float4 *bodies = nullptr;    // [n] sorted {x,y,z,m}
    int *left = nullptr, *right = nullptr;   // [n] binary children
    int *plen = nullptr;         // [n] common-prefix length of internal node
    int *oct = nullptr;          // [n] 1 if the binary node is an octree node
    int *cnt = nullptr;          // [n] octree child count (0: not an octree node)
    int *first = nullptr;        // [n] first child record of octree node
    int *alloc = nullptr;        // [kMaxSeg] record allocators per disk
    int *maxlev = nullptr;       // [kMaxSeg] deepest octree level per disk
    float4 *bin_com = nullptr;   // [n] {com, mass} of binary nodes
    float4 *bin_cell = nullptr;  // [n] {cell centre, half size} of binary nodes
    float4 *rec_a = nullptr;     // walk records {com, mass}
    int4 *rec_b = nullptr;       // walk records {h, h^2, first, count}
    float4 *rec_c = nullptr;     // walk records {cell centre | body, h}
    int n_rec = 0;
    float *partial = nullptr;    // [blocks * 6] bbox partials
    Root *roots = nullptr;       // [kMaxSeg]
    float *dv = nullptr;         // [3 n] collision velocity deltas
    int *err = nullptr;          // [4] error flags / counters
    int *labels = nullptr;       // [n] FoF scratch
};

constexpr int kPasses = 8;       // 8-bit digits over the 63-bit key
constexpr int kSweepItems = 4;
constexpr int kSweepTile = kBlock * kSweepItems;
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;
// tickets[] slots
constexpr int kTicketPass = 0;   // + pass (9 slots incl. the disk pass)
constexpr int kTicketScan = 10;
constexpr int kBlockCtr = 11;    // last-block counter of the bbox reduction


// ---- integrator (reference-identical float ops) --------------------------
void launch_half_kick(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                      const float *az, float dt_half, int n, hipStream_t st);
void launch_drift(float *x, float *y, float *z, const float *vx, const float *vy,
                  const float *vz, float dt, int n, hipStream_t st);
void launch_reset_acc(float *ax, float *ay, float *az, int n, hipStream_t st);
void launch_gas_drag(const float *x, const float *z, float *vx, float *vy, float *vz,
                     float decay, float star_mass, int n, hipStream_t st);
void launch_perturber(float *ax, float *ay, float *az, const float *x, const float *y,
                      const float *z, float px, float py, float pz, float pmass, float soft,
                      int n, hipStream_t st);

// ---- fused step kernels ---------------------------------------------------
// v += a*dt_half; x += v*dt; then per-block bbox partials of the new x.
void launch_kick_drift_bbox(const SegTable &T, float *x, float *y, float *z, float *vx,
                            float *vy, float *vz, const float *ax, const float *ay,
                            const float *az, float dt_half, float dt, Bufs &B, hipStream_t st);
// v += dv (if dv); v += a*dt_half; gas drag (if decay < 1 or has_drag).
void launch_kick_drag(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                      const float *az, const float *dv, const float *x, const float *z,
                      float dt_half, int has_drag, float decay, float star_mass, int n,
                      hipStream_t st);

// ---- tree -----------------------------------------------------------------
void launch_bbox(const SegTable &T, const float *x, const float *y, const float *z, Bufs &B,
                 hipStream_t st);
void launch_root(const SegTable &T, Bufs &B, hipStream_t st);
void launch_build(const SegTable &T, const float *x, const float *y, const float *z,
                  const float *mass, Bufs &B, hipStream_t st);
void launch_summarize(const SegTable &T, Bufs &B, hipStream_t st);
void launch_fill_int(int *p, int v, int n, hipStream_t st);
void launch_fill_float(float *p, float v, int n, hipStream_t st);

// ---- walks ----------------------------------------------------------------
struct WalkArgs {
    float theta2, eps2, softening, star_mass;
    float *ax, *ay, *az;          // original order, written with '='
    float *sax, *say, *saz;       // optional self-gravity outputs
};
void launch_walk(const SegTable &T, const DiskTable &D, const WalkArgs &W, Bufs &B,
                 hipStream_t st);
// Per-body accepted / opened record counts of the gravity walk.
void launch_walk_counts(const SegTable &T, float theta2, int *n_acc, int *n_open, Bufs &B,
                        hipStream_t st);
// Collision deltas into B.dv (original order).  r_coll per disk from D.
void launch_collide(const SegTable &T, const DiskTable &D, const float *vx, const float *vy,
                    const float *vz, float restitution, Bufs &B, hipStream_t st);
// FoF labels (original order) with link length ll for every disk.
void launch_fof(const SegTable &T, float ll, int *labels, Bufs &B, hipStream_t st);

}  // namespace nbx
