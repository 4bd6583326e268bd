// nbx_launch.hpp -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include "nbx_internal.hpp"

namespace nbx {

// N doubles (moment vectors of the prefix scans).
template <int N>
struct DVec {
    double v[N];
};

// Device scratch owned by one nbx_tree.  Sizes are for n bodies over all
// disks; see nbx_internal.hpp for the layout.
struct Bufs {
    int n = 0;
    int n_sweep_tiles = 0;       // radix-sort tiles (kSweepTile keys each)
    int n_scan_tiles = 0;        // prefix-scan tiles (kScanTile bodies each)
    uint64_t *keys[2] = {};      // double-buffered sort keys
    int *vals[2] = {};           // double-buffered body indices
    int *ghist = nullptr;        // [kHistInts] global digit histograms of all passes
    int radix_bits = 0;          // 0: splitter-bucket sort; 8, 9 or 11: radix digit width
    uint64_t *spk = nullptr;     // [buckets] lower splitter key of every bucket (first unused)
    int *spi = nullptr;          // [buckets] body index of that splitter
    int *bm = nullptr;           // [moff] packed {offset << 16 | count} per (bucket, scatter block)
    float4 *spay = nullptr;      // [n] {x, y, z, m} staged beside keys[1] / vals[1]
    bool split_ready = false;    // spk / spi hold the quantiles of the last splitter sort
    // launch order of the stepping walk's waves, by step parity (2 x lpt_stride)
    int *lpt_order = nullptr;    // [2][lpt_stride] wave of each launch slot
    int *lpt_cost = nullptr;     // [2][lpt_stride] work of each wave in its last walk
    int lpt_stride = 0;          // waves of the narrowest walk shape (8 targets)
    int *lpt_hist = nullptr;     // [parts][kLptBins] cost histogram of every part
    unsigned *status[2] = {};    // [n_sweep_tiles * 2048] look-back words per pass
    int *tickets = nullptr;      // [32] tile tickets (per pass / scan) + counters
    DVec<4> *prefix = nullptr;   // [n + 1] exclusive prefix of {m, m x, m y, m z}
    // scan look-back words (one allocation, reset to kScanEmpty by the key
    // pass): tile aggregates and inclusive tile prefixes of both scans
    double *scanw = nullptr;     // [kScanWords * n_scan_tiles]
    DVec<4> *pagg = nullptr;     // [n_scan_tiles] tile aggregates
    DVec<4> *pinc = nullptr;     // [n_scan_tiles] sums of blocks of 64 tiles (at the first)
    int quad = 0;                // 1: quadrupole moments are built and used
    DVec<6> *prefix2 = nullptr;  // [n + 1] exclusive prefix of the second moments
    DVec<6> *pagg2 = nullptr, *pinc2 = nullptr;
    float4 *rec_q = nullptr;     // [2 n_rec] quadrupole of walk records (0 for leaves)
    float4 *bodies = nullptr;    // [n] sorted {x,y,z,m}
    int *left = nullptr, *right = nullptr;   // [n] binary children
    int *plen = nullptr;         // [n] common-prefix length of internal node
    int *oct = nullptr;          // [n] 1 if the binary node is an octree node
    int *cnt = nullptr;          // [n] octree child count (0: not an octree node)
    int *first = nullptr;        // [n] first child record of octree node
    int *alloc = nullptr;        // [kMaxSeg] child records per disk (statistics)
    int *maxlev = nullptr;       // [kMaxSeg] deepest octree level per disk
    float4 *bin_cell = nullptr;  // [n] {cell centre, half size} of binary nodes
    int2 *brange = nullptr;      // [n] sorted-position range [lo, hi] of binary nodes
    float4 *rec_a = nullptr;     // walk records {com, mass}
    int4 *rec_b = nullptr;       // walk records {code, h^2, first, count}
    float4 *rec_c = nullptr;     // walk records {cell centre | body, h}
    int n_rec = 0;
    float *partial = nullptr;    // [blocks * 6] per-block bbox min / max
    Root *roots = nullptr;       // [kMaxSeg]
    float *dv = nullptr;         // [3 n] collision velocity deltas
    int *err = nullptr;          // [4] error flags / counters
    int *labels = nullptr;       // [n] FoF scratch
    int cur = 0;                 // buffer of keys[] / vals[] holding the sorted order
    float *lo = nullptr;         // [6 n] low-order parts of x,y,z,vx,vy,vz (compensated mode)
    double *s64 = nullptr;       // [9 n] FP64 state x,y,z,vx,vy,vz,ax,ay,az (FP64 step)
    double4 *rec64 = nullptr;    // [n_rec] FP64 {com, mass} of walk records (FP64 step)
    double *diag = nullptr;      // [8] FP64 diagnostic sums
    DiskDev *dtab = nullptr;     // [dtab_cap * kMaxSeg] per-step disk values
    int dtab_cap = 0;
    int *step_ctr = nullptr;     // [1] step index into dtab, advanced by k_kick_drag
    // reverse mode (allocated on demand)
    int2 *rec_r = nullptr;       // [n_rec] sorted-position range of walk records
    int *opar = nullptr;         // [n] octree parent (binary index) of octree nodes, -1: root
    int *opar_leaf = nullptr;    // [n] octree parent of every sorted body's leaf record
    int *kids = nullptr;         // [8 n] children codes of every octree node (binary index)
    double *adj_g = nullptr;     // [3 n_rec] per-record sums of J(d) abar over accepting targets
    double *adj_d = nullptr;     // [3 (n + 1)] range-add differences over sorted positions
    double *adj_blk = nullptr;   // [3 blocks] block totals of the difference scan
    double *adj_p = nullptr;     // [4] perturber parameter sums (mass, radius)
};

constexpr int kScanWords = 4 + 4 + 6 + 6;
// bit pattern of a look-back word not yet written (a NaN no sum produces)
constexpr unsigned long long kScanEmpty = 0x7ff8dead0badf00dull;
constexpr int kHistInts = 6 * 2048;  // histograms of the widest setting (6 x 11-bit)
constexpr int kSweepItems = 4;
constexpr int kSweepTile = kBlock * kSweepItems;
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;
// tickets[] slots
constexpr int kTicketPass = 0;   // + pass (9 slots incl. the disk pass)
constexpr int kTicketScan = 10;
constexpr int kTicketScan2 = 12;


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
// v += a*dt_half; x += v*dt; then the per-block boxes of the new x.
void launch_kick_drift_bbox(const SegTable &T, float *x, float *y, float *z, float *vx,
                            float *vy, float *vz, const float *ax, const float *ay,
                            const float *az, float dt_half, float dt, Bufs &B, hipStream_t st,
                            const LptDev &lpt = LptDev{});
// v += dv (if dv); v += a*dt_half; gas drag (if decay < 1 or has_drag).
// step_ctr (optional) is advanced by one after the kernel.
void launch_kick_drag(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                      const float *az, const float *dv, const float *x, const float *z,
                      float dt_half, int has_drag, float decay, float star_mass, int n,
                      int *step_ctr, hipStream_t st);

// The second half-kick with the collision deltas and the drag of one step
// followed by the first half-kick, drift and boxes of the next (the same
// acceleration serves both kicks); step_ctr (optional) advances by one.
void launch_drag_kick_drift(const SegTable &T, float *x, float *y, float *z, float *vx,
                            float *vy, float *vz, const float *ax, const float *ay,
                            const float *az, const float *dv, float dt_half, float dt,
                            int has_drag, float decay, float star_mass, int *step_ctr, Bufs &B,
                            hipStream_t st);

// Compensated variants: the float state holds round(X), lo[] holds
// X - round(X); updates are evaluated in FP64 on X = hi + lo.
void launch_kick_drift_bbox_c(const SegTable &T, float *x, float *y, float *z, float *vx,
                              float *vy, float *vz, const float *ax, const float *ay,
                              const float *az, float dt_half, float dt, Bufs &B, hipStream_t st,
                              const LptDev &lpt = LptDev{});
void launch_kick_drag_c(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                        const float *az, const float *dv, const float *x, const float *z,
                        float dt_half, int has_drag, double decay, float star_mass, int n,
                        int *step_ctr, Bufs &B, hipStream_t st);
// FP64 sums: out[0] kinetic, [1] star potential, [2] L_x, [3] L_y, [4] L_z,
// [5] self potential (pairwise, softened; only if with_self), [6] total mass.
// The state is B.s64 when s64 is given, else the float arrays (+ lo).
void launch_diagnostics(const float *x, const float *y, const float *z, const float *vx,
                        const float *vy, const float *vz, const float *m, const float *lo,
                        const double *s64, int n, double star_mass, double soft, int with_self,
                        Bufs &B, hipStream_t st);

// FP64 step: B.s64 = [x | y | z | vx | vy | vz | ax | ay | az] (9 n doubles,
// body order); the float arrays receive the rounded values after each pass.
void launch_load64(const float *x, const float *y, const float *z, const float *vx,
                   const float *vy, const float *vz, double *s64, int n, hipStream_t st);
// out = [x | y | z | vx | vy | vz | ax | ay | az] as doubles from s64,
// hi + lo, or the floats.
void launch_state64(const float *x, const float *y, const float *z, const float *vx,
                    const float *vy, const float *vz, const float *ax, const float *ay,
                    const float *az, const float *lo, const double *s64, int n, double *out,
                    hipStream_t st);
void launch_kick_drift_bbox_d(const SegTable &T, float *x, float *y, float *z, float *vx,
                              float *vy, float *vz, double dt_half, double dt, Bufs &B,
                              hipStream_t st);
// v += a dt_half; drag (if has_drag); advances step_ctr (if given).
void launch_kick_drag_d(float *vx, float *vy, float *vz, double dt_half, int has_drag,
                        double decay, double star_mass, int n, int *step_ctr, Bufs &B,
                        hipStream_t st);
// Node moments in FP64 (B.rec64 from the FP64 prefix and the state) and the
// FP64 walk: acceleration of every body into B.s64 (and rounded into ax..az).
void launch_rec64(const SegTable &T, Bufs &B, hipStream_t st);
void launch_walk64(const SegTable &T, const DiskDev *dtab, const int *step_ctr, double theta2,
                   double eps2, double star_mass, float *ax, float *ay, float *az, Bufs &B,
                   hipStream_t st);

// ---- tree -----------------------------------------------------------------
void launch_bbox(const SegTable &T, const float *x, const float *y, const float *z, Bufs &B,
                 hipStream_t st);
void launch_root(const SegTable &T, Bufs &B, hipStream_t st);
// resample = false lets the splitter sort reuse the splitters of the previous
// build (the step loop); every other build samples them afresh.  lpt: the
// placement phase of the next walk's wave order (nw = 0: none).  s64: the
// FP64 state ([x | y | z | ...] in body order) whose positions the node
// moments are summed from (the FP64 step); null: the float bodies.
void launch_build(const SegTable &T, const float *x, const float *y, const float *z,
                  const float *mass, Bufs &B, hipStream_t st, bool resample = true,
                  const LptDev &lpt = LptDev{}, const double *s64 = nullptr);
void launch_summarize(const SegTable &T, Bufs &B, hipStream_t st);
// Child records per disk (the root not included) into B.alloc, deepest
// octree level into B.maxlev (statistics only).
void launch_tree_stats(const SegTable &T, Bufs &B, hipStream_t st);
void launch_fill_int(int *p, int v, int n, hipStream_t st);
void launch_fill_float(float *p, float v, int n, hipStream_t st);

// ---- walks ----------------------------------------------------------------
struct WalkArgs {
    float theta2, eps2, softening, star_mass;
    int quad;                     // 1: add the quadrupole of accepted internal records
    int targets;                  // targets per wavefront: 64, 32, 16 or 8
    float *ax, *ay, *az;          // original order, written with '='
    float *sax, *say, *saz;       // optional self-gravity outputs
    int lpt = 0;                  // 1: launch waves longest-first (needs step_ctr)
    int xcd = 0;                  // 1: contiguous wave ranges per XCD (overrides lpt)
};
// Perturber values per disk come from dtab[*step_ctr][disk] when dtab is
// non-null (graph-replayable), otherwise no perturber.
void launch_walk(const SegTable &T, const DiskDev *dtab, const int *step_ctr, const WalkArgs &W,
                 Bufs &B, hipStream_t st);
// Per-body accepted / opened record counts of the gravity walk with the given
// targets per wave; wstat (optional, 3 ints per wave): records popped,
// batches, deepest stack of every wave.
void launch_walk_counts(const SegTable &T, float theta2, int targets, int *n_acc, int *n_open,
                        int *wstat, Bufs &B, hipStream_t st);
int walk_waves(const SegTable &T, int targets);
// Device view of the wave-order sort for a walk of nw waves; the reset puts
// the identity in both orders (when the walk shape changes).
LptDev lpt_dev(Bufs &B, int nw, const int *step_ctr);
void reset_lpt(Bufs &B, hipStream_t st);
// Achievable FP32 TFLOP/s (kind 0: FMA chains, 1: the walk's record mix);
// scratch: a few floats of device memory; -1 on failure.
double probe_fp32(int kind, float *scratch, hipStream_t st);
// The walk and the collision pass of launch_collide as one launch.
void launch_walk_collide(const SegTable &T, const DiskDev *dtab, const int *step_ctr,
                         const WalkArgs &W, const DiskTable &D, const float *vx, const float *vy,
                         const float *vz, float restitution, Bufs &B, hipStream_t st);
// Collision deltas into B.dv (original order).  r_coll per disk from D.
void launch_collide(const SegTable &T, const DiskTable &D, const float *vx, const float *vy,
                    const float *vz, float restitution, Bufs &B, hipStream_t st);
// FoF labels (original order), disk s linked at LT.ll[s].
void launch_fof(const SegTable &T, const LinkTable &LT, int *labels, Bufs &B, hipStream_t st);

// ---- reverse mode (collision-free step, monopole walk) ---------------------
struct AdjWalkArgs {
    float theta2, eps2, softening, star_mass;
    int targets;                  // targets per wavefront: 64, 32, 16 or 8
    float px, py, pz, pmass;      // perturber of this force evaluation (pmass 0: none)
    float drx, dry, drz;          // dX_p / dr_p
    const double *ab;             // [3 n] adjoint of the acceleration (original order)
    double *xb;                   // [3 n] adjoint of the positions, accumulated
};
// xb += (dA/dx)^T ab over the current tree (built with B.brange / B.rec_r);
// B.adj_p[0..1] = dL/dm_p, dL/dr_p of this evaluation.
void launch_adj_force(const SegTable &T, const AdjWalkArgs &A, Bufs &B, hipStream_t st);
int adj_scan_blocks(int n);
void launch_adj_kick_drag(const float *x1, const float *z1, const float *vx0, const float *vy0,
                          const float *vz0, const float *ax0, const float *ay0, const float *az0,
                          const float *ax1, const float *ay1, const float *az1, float dt_half,
                          int has_drag, float decay, float star_mass, double *xb, double *vb,
                          double *ab, int n, hipStream_t st);
void launch_adj_drift_kick(const double *xb, double *vb, double *ab, float dt, float dt_half,
                           int n, hipStream_t st);

}  // namespace nbx
