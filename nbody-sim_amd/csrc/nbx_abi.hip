// nbx_abi.hip -- the extern "C" entry points of include/nbx.h.
//
// Host-side sequencing only: argument checks, device allocation, launch
// order of the gfx950 kernels, error propagation (a failed HIP call or a
// kernel error flag becomes a negative return code and a message), and the
// fused step loop of main.cu:101-140.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nbx.h"
#include "nbx_launch.hpp"

using namespace nbx;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(NBX_E_HIP, "HIP error in %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

enum Stage { kNone = 0, kBBox = 1, kBuilt = 2, kSummarized = 3 };

constexpr int kDefaultTargets = 16;
const int kShapes[4] = {64, 32, 16, 8};

}  // namespace

namespace nbx {
// Shared with the host-side run support (nbx_io.cpp).
void set_last_error(const char *msg) { g_err = msg; }
}  // namespace nbx

struct nbx_tree {
    SegTable T{};
    Bufs B{};
    hipStream_t stream = nullptr;  // stream every entry point enqueues on
    hipStream_t owned = nullptr;   // the stream the tree created (destroyed with it)
    bool async = false;
    int stage = kNone;
    int *h_err = nullptr;  // pinned copy of B.err
    bool timing = false;
    std::vector<hipEvent_t> ev;  // timing events: 6 per step
    float phase_ms[5] = {0, 0, 0, 0, 0};
    int phase_n = 0;
    // per-step disk table staging (pinned) and the event that guards it
    DiskDev *h_dtab = nullptr;
    hipEvent_t dtab_done = nullptr;
    // captured single step: replayed once per step with the device step counter
    hipGraphExec_t graph = nullptr;
    // float steps with the next step's first half-kick fused into the end
    // of the step: `graph` is every step but the last, `graph_last` the last
    hipGraphExec_t graph_last = nullptr;
    bool fused = false;
    std::vector<unsigned char> graph_key;
    // NBX_PRECISION_*: float step, FP64-compensated KDK (B.lo holds the low
    // parts) or FP64 step (B.s64 holds the state once s64_ready)
    int precision = NBX_PRECISION_FLOAT;
    bool s64_ready = false;
    int walk_tg = kDefaultTargets;  // gravity-walk targets per wavefront
    float tune_ms[4] = {0, 0, 0, 0};  // autotuner timings for 64, 32, 16, 8 targets
    // launch order of the walk's waves (NBX_WAVES_*): longest first from the
    // last step's costs (the stepping walk), or contiguous Morton ranges per
    // XCD; lpt_tg: the walk shape the longest-first buffers hold
    int wave_order = NBX_WAVES_XCD;
    bool lpt = false;
    int lpt_tg = 0;
    bool adjoint = false;  // builds record sorted ranges; reverse-mode buffers exist
    double *h_adj = nullptr;  // pinned copy of B.adj_p
    int cus = 0;  // compute units of the device the tree was created on
    // orders the streams of a group of trees stepped together
    // (nbx_step_concurrent); never recorded inside a stream capture
    hipEvent_t gord = nullptr;
};

namespace {

const char *kPhases[5] = {"kdk1", "build", "walk", "collide", "kdk2"};

int finish(nbx_tree *t)
{
    HIPCHK(hipGetLastError());
    if (t->async) return NBX_OK;
    HIPCHK(hipMemcpyAsync(t->h_err, t->B.err, 4 * sizeof(int), hipMemcpyDeviceToHost,
                          t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    if (t->h_err[0]) {  // report once, then clear for the next call
        launch_fill_int(t->B.err, 0, 1, t->stream);
        HIPCHK(hipStreamSynchronize(t->stream));
    }
    if (t->h_err[0] & kErrStack)
        return fail(NBX_E_OVERFLOW, "tree walk stack overflow (more than %d pending blocks)",
                    kStackCap);
    if (t->h_err[0] & kErrTree) return fail(NBX_E_STATE, "inconsistent octree (internal error)");
    return NBX_OK;
}

int finish_plain()
{
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return NBX_OK;
}

bool particles_ok(const nbx_particles *p)
{
    return p && p->n > 0 && p->x && p->y && p->z && p->vx && p->vy && p->vz && p->ax && p->ay &&
           p->az && p->mass;
}

template <class T>
int dmalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void **)p, count * sizeof(T)));
    return NBX_OK;
}

// The scratch of the tree build, summarize, walks and searches (every buffer
// the FoF path touches), in one fixed order: take(ptr, count) assigns each.
template <class Take>
void core_layout(const SegTable &T, Bufs &B, Take &&take)
{
    const int n = T.n_total;
    B.n = n;
    B.radix_bits = 0;
    for (int s = 0; s < T.nseg; ++s)
        if (T.off[s + 1] - T.off[s] > kSplitMaxN) B.radix_bits = 8;
    B.n_sweep_tiles = (n + kSweepTile - 1) / kSweepTile;
    B.n_scan_tiles = T.stile[T.nseg];
    B.n_rec = T.noff[T.nseg] + kRecPad;
    take(B.keys[0], n);
    take(B.keys[1], n);
    take(B.vals[0], n);
    take(B.vals[1], n);
    take(B.ghist, kHistInts);
    take(B.spk, T.bkoff[T.nseg]);
    take(B.spi, T.bkoff[T.nseg]);
    take(B.bm, T.moff[T.nseg]);
    take(B.spay, n);
    take(B.tickets, 32);
    take(B.prefix, n + T.nseg);
    // scan look-back words of both prefix scans: 4 + 4 + 6 + 6 doubles per tile
    take(B.scanw, (size_t)kScanWords * B.n_scan_tiles);
    take(B.bodies, n);
    take(B.left, n);
    take(B.right, n);
    take(B.plen, n);
    take(B.oct, n);
    take(B.cnt, n);
    take(B.first, n);
    take(B.alloc, 2 * kMaxSeg);
    take(B.opar, n);
    take(B.opar_leaf, n);
    take(B.kids, 8 * (size_t)n);
    take(B.brange, n);
    take(B.bin_cell, n);
    take(B.rec_a, B.n_rec);
    take(B.rec_b, B.n_rec);
    take(B.rec_c, B.n_rec);
    take(B.partial, 6 * T.blk[T.nseg]);
    take(B.roots, kMaxSeg);
    take(B.err, 4);
    take(B.labels, n);
}

// Views into the look-back words (after the layout assigned scanw).
void scan_views(Bufs &B)
{
    B.pagg = (DVec<4> *)B.scanw;
    B.pinc = B.pagg + B.n_scan_tiles;
    B.pagg2 = (DVec<6> *)(B.scanw + 8 * (size_t)B.n_scan_tiles);
    B.pinc2 = B.pagg2 + B.n_scan_tiles;
    B.maxlev = B.alloc + kMaxSeg;  // adjacent: cleared by one fill
}

int alloc_bufs(nbx_tree *t)
{
    Bufs &B = t->B;
    const int n = t->T.n_total;
    int rc = NBX_OK;
    auto take = [&rc](auto *&ptr, size_t count) {
        if (rc == NBX_OK) rc = dmalloc(&ptr, count);
    };
    core_layout(t->T, B, take);
    if (rc != NBX_OK) return rc;
    scan_views(B);
    B.lpt_stride = walk_waves(t->T, 8);
    take(B.lpt_order, 2 * (size_t)B.lpt_stride);
    take(B.lpt_cost, 2 * (size_t)B.lpt_stride);
    take(B.lpt_hist, (size_t)lpt_parts(B.lpt_stride) * kLptBins);
    take(B.status[0], 2048 * (size_t)B.n_sweep_tiles);
    take(B.status[1], 2048 * (size_t)B.n_sweep_tiles);
    take(B.dv, 3 * (size_t)n);
    take(B.step_ctr, 1);
    if (rc != NBX_OK) return rc;
    HIPCHK(hipMemset(B.rec_a, 0, sizeof(float4) * B.n_rec));
    HIPCHK(hipMemset(B.rec_b, 0, sizeof(int4) * B.n_rec));
    HIPCHK(hipMemset(B.rec_c, 0, sizeof(float4) * B.n_rec));
    HIPCHK(hipMemset(B.err, 0, 4 * sizeof(int)));
    HIPCHK(hipMemset(B.dv, 0, 3 * sizeof(float) * n));
    HIPCHK(hipMemset(B.ghist, 0, kHistInts * sizeof(int)));
    HIPCHK(hipMemset(B.tickets, 0, 32 * sizeof(int)));
    HIPCHK(hipMemset(B.alloc, 0, 2 * kMaxSeg * sizeof(int)));

    HIPCHK(hipHostMalloc((void **)&t->h_err, 4 * sizeof(int)));
    return NBX_OK;
}

void free_bufs(nbx_tree *t)
{
    Bufs &B = t->B;
    void *ptrs[] = {B.keys[0], B.keys[1], B.vals[0], B.vals[1], B.ghist,   B.status[0],
                    B.status[1], B.tickets, B.prefix, B.scanw,
                    B.bodies,  B.left,    B.right,   B.plen,    B.oct,     B.cnt,
                    B.first,   B.alloc,   B.bin_cell, B.rec_a,  B.rec_b,
                    B.rec_c,   B.partial, B.roots,   B.dv,      B.err,     B.labels,
                    B.step_ctr, B.dtab,    B.prefix2,
                    B.rec_q,   B.lo,      B.diag,    B.brange,  B.rec_r,
                    B.adj_g,   B.adj_d,   B.adj_blk, B.adj_p,   B.spk,     B.spi,
                    B.bm,      B.lpt_order, B.lpt_cost, B.opar, B.opar_leaf, B.spay,
                    B.kids,    B.lpt_hist, B.s64,  B.rec64};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (t->h_err) (void)hipHostFree(t->h_err);
    if (t->h_dtab) (void)hipHostFree(t->h_dtab);
    if (t->h_adj) (void)hipHostFree(t->h_adj);
    if (t->dtab_done) (void)hipEventDestroy(t->dtab_done);
    if (t->graph) (void)hipGraphExecDestroy(t->graph);
    if (t->graph_last) (void)hipGraphExecDestroy(t->graph_last);
    t->graph = t->graph_last = nullptr;
    B = Bufs{};
    t->h_err = nullptr;
    t->h_dtab = nullptr;
    t->dtab_done = nullptr;
    t->graph = nullptr;
}

int check_tree_particles(nbx_tree *t, const nbx_particles *p)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (!particles_ok(p)) return fail(NBX_E_ARG, "invalid particle arrays");
    if (p->n != t->T.n_total)
        return fail(NBX_E_ARG, "particle count %d != tree bodies %d", p->n, t->T.n_total);
    return NBX_OK;
}

// What the FP64 step leaves out: the collision pass and quadrupoles.
int fp64_checks(nbx_tree *t, const nbx_step_params *sp)
{
    if (sp->collisions) return fail(NBX_E_ARG, "the FP64 step has no collision pass");
    if (t->B.quad) return fail(NBX_E_ARG, "the FP64 step walks monopoles only");
    return NBX_OK;
}

// Both captured steps (whichever exist), for a change of what they bake in.
hipError_t drop_graphs(nbx_tree *t)
{
    hipError_t e = hipSuccess;
    for (hipGraphExec_t *g : {&t->graph, &t->graph_last}) {
        if (*g) {
            const hipError_t r = hipGraphExecDestroy(*g);
            if (e == hipSuccess) e = r;
            *g = nullptr;
        }
    }
    t->graph_key.clear();
    return e;
}

// Streams and events of a tree (whichever were created), then the tree.
void destroy_handles(nbx_tree *t)
{
    for (hipEvent_t e : t->ev) (void)hipEventDestroy(e);
    t->ev.clear();
    if (t->owned) (void)hipStreamDestroy(t->owned);
    if (t->gord) (void)hipEventDestroy(t->gord);
    delete t;
}

// Per-disk offsets of nseg disks of counts[s] bodies (NBX_OK, or an
// argument error with its message).
int make_segtable(int nseg, const int32_t *counts, SegTable &T)
{
    if (nseg < 1 || nseg > kMaxSeg) return fail(NBX_E_ARG, "disk count %d not in [1,%d]", nseg,
                                               kMaxSeg);
    T = SegTable{};
    T.nseg = nseg;
    long total = 0;
    for (int s = 0; s < nseg; ++s) {
        if (counts[s] <= 0) return fail(NBX_E_ARG, "disk %d has %d bodies", s, counts[s]);
        T.off[s] = (int)total;
        total += counts[s];
    }
    // walk-record indices (8 n + 8 per disk) stay below 2^31
    if (total > (1L << 27)) return fail(NBX_E_ARG, "too many bodies (%ld > 2^27)", total);
    T.off[nseg] = (int)total;
    T.n_total = (int)total;
    for (int s = 0; s < nseg; ++s) {
        const int n = counts[s];
        const int nb = std::min(kMaxBuckets, (n + kBucketMean - 1) / kBucketMean);
        const int nblk = (n + kSplitTile - 1) / kSplitTile;
        T.bkoff[s + 1] = T.bkoff[s] + nb;
        T.sblk[s + 1] = T.sblk[s] + nblk;
        T.moff[s + 1] = T.moff[s] + nb * nblk;
        T.stile[s + 1] = T.stile[s] + (n + kScanTile - 1) / kScanTile;
        T.blk[s + 1] = T.blk[s] + (n + kBlock - 1) / kBlock;
        T.wave[s + 1] = T.wave[s] + (n + kWave - 1) / kWave;
        T.ioff[s + 1] = T.ioff[s] + (n - 1);
        T.noff[s + 1] = T.noff[s] + 8 * n + 8;  // root slot + an 8-record block per node
    }
    for (int s = nseg + 1; s <= kMaxSeg; ++s) {
        T.off[s] = T.off[nseg];
        T.blk[s] = T.blk[nseg];
        T.wave[s] = T.wave[nseg];
        T.ioff[s] = T.ioff[nseg];
        T.noff[s] = T.noff[nseg];
        T.stile[s] = T.stile[nseg];
        T.bkoff[s] = T.bkoff[nseg];
        T.sblk[s] = T.sblk[nseg];
        T.moff[s] = T.moff[nseg];
    }
    return NBX_OK;
}

int create_tree(nbx_tree **out, int nseg, const int32_t *counts)
{
    if (!out) return fail(NBX_E_ARG, "null output pointer");
    *out = nullptr;
    SegTable T;
    int rc0 = make_segtable(nseg, counts, T);
    if (rc0) return rc0;
    nbx_tree *t = new nbx_tree();
    t->T = T;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&t->cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipStreamCreate(&t->owned);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&t->gord, hipEventDisableTiming);
    t->stream = t->owned;
    if (e != hipSuccess) {
        std::string msg = std::string("stream / event setup: ") + hipGetErrorString(e);
        destroy_handles(t);
        return fail(NBX_E_HIP, "%s", msg.c_str());
    }
    int rc = alloc_bufs(t);
    if (rc != NBX_OK) {
        std::string msg = g_err;
        free_bufs(t);
        destroy_handles(t);
        g_err = msg;
        return rc;
    }
    *out = t;
    return NBX_OK;
}

// r_coll of every disk (constant over a run): passed by value to the walks.
DiskTable disk_rcoll(const nbx_tree *t, const nbx_disk_params *dp)
{
    DiskTable D;
    memset(&D, 0, sizeof(D));
    for (int s = 0; s < t->T.nseg; ++s) D.d[s].r_coll = dp ? dp[s].r_coll : 0.f;
    return D;
}

// Per-step perturber positions of steps [step0, step0+n) for every disk,
// from updatePerturber (Perturber.h:60-66) in float on the host, staged in
// pinned memory and copied to B.dtab on the tree's stream; the step counter
// is reset so the walk of step k reads row k.
int upload_steps(nbx_tree *t, const nbx_disk_params *dp, int step0, int n, float dt, bool t0)
{
    Bufs &B = t->B;
    if (n > B.dtab_cap) {
        HIPCHK(hipStreamSynchronize(t->stream));
        if (B.dtab) HIPCHK(hipFree(B.dtab));
        if (t->h_dtab) HIPCHK(hipHostFree(t->h_dtab));
        B.dtab = nullptr;
        t->h_dtab = nullptr;
        const int cap = n < 64 ? 64 : n;
        HIPCHK(hipMalloc((void **)&B.dtab, sizeof(DiskDev) * kMaxSeg * cap));
        HIPCHK(hipHostMalloc((void **)&t->h_dtab, sizeof(DiskDev) * kMaxSeg * cap));
        B.dtab_cap = cap;
        HIPCHK(drop_graphs(t));  // the captured step holds the old table address
    }
    if (!t->dtab_done) HIPCHK(hipEventCreateWithFlags(&t->dtab_done, hipEventDisableTiming));
    HIPCHK(hipEventSynchronize(t->dtab_done));  // the previous upload has left the staging
    memset(t->h_dtab, 0, sizeof(DiskDev) * kMaxSeg * n);
    for (int k = 0; k < n; ++k) {
        const float sim_time = t0 ? 0.0f : (float)(step0 + k + 1) * dt;  // main.cu:110
        for (int s = 0; s < t->T.nseg; ++s) {
            DiskDev &d = t->h_dtab[k * kMaxSeg + s];
            d.r_coll = dp[s].r_coll;
            if (!dp[s].perturber_enabled) continue;
            float phi = dp[s].pert_phase0 + dp[s].pert_omega * sim_time;
            d.px = dp[s].pert_radius_au * cosf(phi);
            d.pz = dp[s].pert_radius_au * sinf(phi);
            d.py = 0.f;
            d.pmass = dp[s].pert_mass_msun;
        }
    }
    HIPCHK(hipMemcpyAsync(B.dtab, t->h_dtab, sizeof(DiskDev) * kMaxSeg * n,
                          hipMemcpyHostToDevice, t->stream));
    HIPCHK(hipEventRecord(t->dtab_done, t->stream));
    launch_fill_int(B.step_ctr, 0, 1, t->stream);
    return NBX_OK;
}

void ev_record(nbx_tree *t, int idx)
{
    if (t->timing) (void)hipEventRecord(t->ev[idx], t->stream);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *nbx_last_error(void) { return g_err.c_str(); }

int nbx_particles_alloc(nbx_particles *p, int32_t n)
{
    if (!p || n <= 0) return fail(NBX_E_ARG, "nbx_particles_alloc: invalid arguments");
    memset(p, 0, sizeof(*p));
    float **f[10] = {&p->x, &p->y, &p->z, &p->vx, &p->vy, &p->vz, &p->ax, &p->ay, &p->az,
                     &p->mass};
    for (int k = 0; k < 10; ++k) {
        HIPCHK(hipMalloc((void **)f[k], sizeof(float) * n));
        HIPCHK(hipMemset(*f[k], 0, sizeof(float) * n));
    }
    HIPCHK(hipMalloc((void **)&p->id, sizeof(int32_t) * n));
    HIPCHK(hipMemset(p->id, 0, sizeof(int32_t) * n));
    p->n = n;
    return NBX_OK;
}

int nbx_particles_free(nbx_particles *p)
{
    if (!p) return fail(NBX_E_ARG, "null particles");
    void *f[11] = {p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay, p->az, p->mass, p->id};
    for (void *q : f)
        if (q) (void)hipFree(q);
    memset(p, 0, sizeof(*p));
    return NBX_OK;
}

static int copy_particles(const nbx_particles *src, nbx_particles *dst, hipMemcpyKind kind,
                          const char *what)
{
    if (!src || !dst) return fail(NBX_E_ARG, "%s: null particles", what);
    if (src->n != dst->n)
        return fail(NBX_E_ARG, "%s: source n=%d != destination n=%d", what, src->n, dst->n);
    const size_t fb = sizeof(float) * src->n;
    const float *s[10] = {src->x, src->y, src->z, src->vx, src->vy, src->vz, src->ax, src->ay,
                          src->az, src->mass};
    float *d[10] = {dst->x, dst->y, dst->z, dst->vx, dst->vy, dst->vz, dst->ax, dst->ay,
                    dst->az, dst->mass};
    for (int k = 0; k < 10; ++k) HIPCHK(hipMemcpy(d[k], s[k], fb, kind));
    if (src->id && dst->id) HIPCHK(hipMemcpy(dst->id, src->id, sizeof(int32_t) * src->n, kind));
    return NBX_OK;
}

int nbx_copy_to_device(const nbx_particles *host, nbx_particles *dev)
{
    return copy_particles(host, dev, hipMemcpyHostToDevice, "copyToGPU");
}

int nbx_copy_to_host(const nbx_particles *dev, nbx_particles *host)
{
    return copy_particles(dev, host, hipMemcpyDeviceToHost, "copyToCPU");
}

int nbx_tree_create(nbx_tree **t, int32_t n_bodies)
{
    return create_tree(t, 1, &n_bodies);
}

int nbx_tree_create_batch(nbx_tree **t, int32_t n_seg, const int32_t *seg_counts)
{
    if (!seg_counts) return fail(NBX_E_ARG, "null disk counts");
    return create_tree(t, n_seg, seg_counts);
}

int nbx_tree_destroy(nbx_tree *t)
{
    if (!t) return NBX_OK;
    (void)hipStreamSynchronize(t->stream);
    free_bufs(t);
    destroy_handles(t);
    return NBX_OK;
}

int nbx_tree_reset(nbx_tree *t)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    launch_fill_float(t->B.dv, 0.f, 3 * t->T.n_total, t->stream);
    launch_fill_int(t->B.err, 0, 4, t->stream);
    t->stage = kNone;
    return finish(t);
}

int nbx_tree_set_stream(nbx_tree *t, void *stream)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    HIPCHK(hipStreamSynchronize(t->stream));
    // the tree's own stream stays alive (and owned) while a caller's stream
    // is in use; only the owned one is destroyed with the tree
    t->stream = stream ? (hipStream_t)stream : t->owned;
    return NBX_OK;
}

int nbx_tree_set_multipole(nbx_tree *t, int order)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (order != 1 && order != 2) return fail(NBX_E_ARG, "multipole order %d not in {1, 2}", order);
    Bufs &B = t->B;
    if (order == 2 && !B.prefix2) {
        const int n = t->T.n_total;
        int rc;
        if ((rc = dmalloc(&B.prefix2, (size_t)n + t->T.nseg)) != NBX_OK) return rc;
        if ((rc = dmalloc(&B.rec_q, (size_t)2 * B.n_rec)) != NBX_OK) return rc;
        HIPCHK(hipMemset(B.rec_q, 0, sizeof(float4) * 2 * B.n_rec));
    }
    if (B.quad != (order == 2)) {
        B.quad = order == 2;
        t->stage = kNone;  // records must be rebuilt with (or without) moments
        HIPCHK(drop_graphs(t));
    }
    return NBX_OK;
}

int nbx_tree_set_precision(nbx_tree *t, int mode)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (mode != NBX_PRECISION_FLOAT && mode != NBX_PRECISION_COMPENSATED &&
        mode != NBX_PRECISION_FP64)
        return fail(NBX_E_ARG, "precision %d not in {0, 1, 2}", mode);
    Bufs &B = t->B;
    const int n = t->T.n_total;
    if (mode == NBX_PRECISION_COMPENSATED) {
        if (!B.lo) {
            int rc = dmalloc(&B.lo, (size_t)6 * n);
            if (rc) return rc;
        }
        HIPCHK(hipStreamSynchronize(t->stream));
        HIPCHK(hipMemset(B.lo, 0, sizeof(float) * 6 * n));  // state = its float values
    }
    if (mode == NBX_PRECISION_FP64) {
        // stack entries of the FP64 walk pack a record index above 4 bits
        if (t->T.noff[t->T.nseg] >= (1 << 27))
            return fail(NBX_E_ARG, "the FP64 step takes at most 2^24 bodies");
        int rc;
        if ((!B.s64 && (rc = dmalloc(&B.s64, (size_t)9 * n))) ||
            (!B.rec64 && (rc = dmalloc(&B.rec64, (size_t)B.n_rec))))
            return rc;
        t->stage = kNone;  // the FP64 moments come with the next build
        t->s64_ready = false;  // loaded from the float state by nbx_init_forces
    }
    if (t->precision != mode) HIPCHK(drop_graphs(t));
    t->precision = mode;
    return NBX_OK;
}

int nbx_state64(nbx_tree *t, const nbx_particles *p, double *dev_out)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (!dev_out) return fail(NBX_E_ARG, "null output");
    const bool s64 = t->precision == NBX_PRECISION_FP64 && t->s64_ready;
    launch_state64(p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay, p->az,
                   t->precision == NBX_PRECISION_COMPENSATED ? t->B.lo : nullptr,
                   s64 ? t->B.s64 : nullptr, p->n, dev_out, t->stream);
    return finish(t);
}

int nbx_diagnostics(nbx_tree *t, const nbx_particles *p, float star_mass, float softening,
                    int with_self, double *out)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (!out) return fail(NBX_E_ARG, "null output");
    if (!t->B.diag) {
        rc = dmalloc(&t->B.diag, 8);
        if (rc) return rc;
    }
    const bool s64 = t->precision == NBX_PRECISION_FP64 && t->s64_ready;
    launch_diagnostics(p->x, p->y, p->z, p->vx, p->vy, p->vz, p->mass,
                       t->precision == NBX_PRECISION_COMPENSATED ? t->B.lo : nullptr,
                       s64 ? t->B.s64 : nullptr, p->n, star_mass, softening, with_self, t->B,
                       t->stream);
    HIPCHK(hipMemcpyAsync(out, t->B.diag, 7 * sizeof(double), hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    return finish(t);
}

int nbx_tree_set_radix_bits(nbx_tree *t, int bits)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (bits != 0 && bits != 8 && bits != 9 && bits != 11)
        return fail(NBX_E_ARG, "radix bits %d not in {0, 8, 9, 11}", bits);
    if (bits == 0)
        for (int s = 0; s < t->T.nseg; ++s)
            if (t->T.off[s + 1] - t->T.off[s] > kSplitMaxN)
                return fail(NBX_E_ARG, "splitter sort takes disks of at most %d bodies", kSplitMaxN);
    t->B.radix_bits = bits;
    t->stage = kNone;
    HIPCHK(drop_graphs(t));
    return NBX_OK;
}

int nbx_tree_set_wave_order(nbx_tree *t, int order)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (order != NBX_WAVES_LAUNCH && order != NBX_WAVES_LONGEST_FIRST && order != NBX_WAVES_XCD)
        return fail(NBX_E_ARG, "wave order %d not in {0, 1, 2}", order);
    t->wave_order = order;
    t->lpt = order == NBX_WAVES_LONGEST_FIRST;
    t->lpt_tg = 0;  // longest-first buffers restart from the identity when used again
    HIPCHK(drop_graphs(t));
    return NBX_OK;
}

int nbx_tree_set_walk_shape(nbx_tree *t, int targets_per_wave)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    bool ok = false;
    for (int k : kShapes) ok |= k == targets_per_wave;
    if (!ok) return fail(NBX_E_ARG, "targets per wave %d not in {64, 32, 16, 8}", targets_per_wave);
    t->walk_tg = targets_per_wave;
    return NBX_OK;
}

int nbx_tree_get_walk_shape(nbx_tree *t, int32_t *targets_per_wave)
{
    if (!t || !targets_per_wave) return fail(NBX_E_ARG, "null argument");
    *targets_per_wave = t->walk_tg;
    return NBX_OK;
}

int nbx_tune_walk(nbx_tree *t, float theta, float softening, float star_mass, float *ms_out)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "walk tuning before summarize");
    const int n = t->T.n_total;
    float *sc = t->B.dv;  // scratch outputs: the tuner never touches the state
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int best = t->walk_tg;
    float best_ms = 1e30f;
    for (int q = 0; q < 4; ++q) {
        WalkArgs W{theta * theta, softening * softening, softening, star_mass, t->B.quad,
                   kShapes[q], sc, sc + n, sc + 2 * n, nullptr, nullptr, nullptr, 0,
                   t->wave_order == NBX_WAVES_XCD ? 1 : 0};
        launch_walk(t->T, nullptr, nullptr, W, t->B, t->stream);  // warm-up
        HIPCHK(hipEventRecord(e0, t->stream));
        for (int rep = 0; rep < 3; ++rep) launch_walk(t->T, nullptr, nullptr, W, t->B, t->stream);
        HIPCHK(hipEventRecord(e1, t->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        t->tune_ms[q] = ms / 3.f;
        if (ms_out) ms_out[q] = ms / 3.f;
        if (ms < best_ms) {
            best_ms = ms;
            best = kShapes[q];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    t->walk_tg = best;
    return finish(t);
}

int nbx_tree_set_async(nbx_tree *t, int async)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    t->async = async != 0;
    return NBX_OK;
}

int nbx_tree_sync(nbx_tree *t)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    const bool a = t->async;
    t->async = false;
    int rc = finish(t);
    t->async = a;
    return rc;
}

int nbx_bbox(nbx_tree *t, const nbx_particles *p)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    launch_bbox(t->T, p->x, p->y, p->z, t->B, t->stream);
    launch_root(t->T, t->B, t->stream);
    t->stage = kBBox;
    return finish(t);
}

int nbx_build(nbx_tree *t, const nbx_particles *p)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    // the build consumes the counters and root cube the bbox pass prepares
    if (t->stage != kBBox) return fail(NBX_E_STATE, "launchBuildKernel must follow launchBBoxKernel");
    launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream);
    t->stage = kBuilt;
    return finish(t);
}

int nbx_summarize(nbx_tree *t)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (t->stage < kBuilt) return fail(NBX_E_STATE, "launchSummarizeKernel before build");
    if (t->stage == kSummarized) return finish(t);  // the tree is summarized already
    launch_summarize(t->T, t->B, t->stream);
    t->stage = kSummarized;
    return finish(t);
}

int nbx_sort(nbx_tree *t)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (t->stage < kBuilt) return fail(NBX_E_STATE, "launchSortKernel before build");
    return finish(t);
}

int nbx_sort_idx(nbx_tree *t, int32_t *dev_out)
{
    if (!t || !dev_out) return fail(NBX_E_ARG, "null argument");
    if (t->stage < kBuilt) return fail(NBX_E_STATE, "sort_idx before build");
    HIPCHK(hipMemcpyAsync(dev_out, t->B.vals[t->B.cur], sizeof(int32_t) * t->T.n_total,
                          hipMemcpyDeviceToDevice, t->stream));
    return finish(t);
}

int nbx_sorted_keys(nbx_tree *t, uint64_t *dev_out)
{
    if (!t || !dev_out) return fail(NBX_E_ARG, "null argument");
    if (t->stage < kBuilt) return fail(NBX_E_STATE, "keys before build");
    HIPCHK(hipMemcpyAsync(dev_out, t->B.keys[t->B.cur], sizeof(uint64_t) * t->T.n_total,
                          hipMemcpyDeviceToDevice, t->stream));
    return finish(t);
}

int nbx_force_ex(nbx_tree *t, nbx_particles *p, float theta, float softening, float star_mass,
                 float *self_ax, float *self_ay, float *self_az)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "launchForceKernel before summarize");
    if (!(theta > 0.f)) return fail(NBX_E_ARG, "theta must be positive");
    WalkArgs W{theta * theta, softening * softening, softening, star_mass, t->B.quad, t->walk_tg,
               p->ax, p->ay,
               p->az, self_ax, self_ay, self_az, 0, t->wave_order == NBX_WAVES_XCD ? 1 : 0};
    launch_walk(t->T, nullptr, nullptr, W, t->B, t->stream);
    return finish(t);
}

int nbx_force(nbx_tree *t, nbx_particles *p, float theta, float softening, float star_mass)
{
    return nbx_force_ex(t, p, theta, softening, star_mass, nullptr, nullptr, nullptr);
}

int nbx_walk_counts(nbx_tree *t, float theta, int32_t *dev_n_acc, int32_t *dev_n_open)
{
    if (!t || !dev_n_acc || !dev_n_open) return fail(NBX_E_ARG, "null argument");
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "walk counts before summarize");
    launch_walk_counts(t->T, theta * theta, 64, dev_n_acc, dev_n_open, nullptr, t->B, t->stream);
    return finish(t);
}

}  // extern "C"

namespace {

// Counting walk with the current shape: per-body accepted / opened counts
// (host, 2 n ints) and per-wave {records, batches, deepest stack} (host,
// 3 ints per wave).
int walk_profile(nbx_tree *t, float theta, std::vector<int> &counts, std::vector<int> &waves)
{
    const int n = t->T.n_total;
    const int nw = walk_waves(t->T, t->walk_tg);
    int *acc = nullptr, *ws = nullptr;
    int rc = dmalloc(&acc, (size_t)2 * n);
    if (rc) return rc;
    if ((rc = dmalloc(&ws, (size_t)3 * nw)) != NBX_OK) {
        (void)hipFree(acc);
        return rc;
    }
    launch_walk_counts(t->T, theta * theta, t->walk_tg, acc, acc + n, ws, t->B, t->stream);
    counts.resize(2 * (size_t)n);
    waves.resize(3 * (size_t)nw);
    hipError_t e = hipMemcpyAsync(counts.data(), acc, sizeof(int) * 2 * n,
                                  hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(waves.data(), ws, sizeof(int) * 3 * nw, hipMemcpyDeviceToHost,
                           t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    (void)hipFree(acc);
    (void)hipFree(ws);
    if (e != hipSuccess) return fail(NBX_E_HIP, "walk profile: %s", hipGetErrorString(e));
    return NBX_OK;
}

}  // namespace

extern "C" {

int nbx_walk_stats(nbx_tree *t, float theta, double *out)
{
    if (!t || !out) return fail(NBX_E_ARG, "null argument");
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "walk statistics before summarize");
    std::vector<int> h, w;
    int rc = walk_profile(t, theta, h, w);
    if (rc) return rc;
    const int n = t->T.n_total, nw = (int)(w.size() / 3);
    double sa = 0, so = 0, sp = 0;
    int depth = 0;
    for (int i = 0; i < n; ++i) {
        sa += h[i];
        so += h[n + i];
    }
    for (int k = 0; k < nw; ++k) {
        sp += w[3 * k];
        depth = depth > w[3 * k + 2] ? depth : w[3 * k + 2];
    }
    out[0] = sa / n;
    out[1] = so / n;
    out[2] = sp / nw;
    out[3] = depth;
    return finish(t);
}

int nbx_walk_wave_stats(nbx_tree *t, float theta, int32_t *out, int32_t cap, int32_t *n_waves)
{
    if (!t || !n_waves) return fail(NBX_E_ARG, "null argument");
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "walk statistics before summarize");
    const int nw = walk_waves(t->T, t->walk_tg);
    *n_waves = nw;
    if (!out) return NBX_OK;
    if (cap < nw) return fail(NBX_E_ARG, "output holds %d waves, the walk has %d", cap, nw);
    std::vector<int> h, w;
    int rc = walk_profile(t, theta, h, w);
    if (rc) return rc;
    memcpy(out, w.data(), sizeof(int) * w.size());
    return finish(t);
}

int nbx_search_stats(nbx_tree *t, int32_t *out, int clear)
{
    if (!t || !out) return fail(NBX_E_ARG, "null argument");
    int h[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(h, t->B.err, sizeof(h), hipMemcpyDeviceToHost, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
    out[0] = h[1];
    out[1] = h[3];
    if (clear) {
        launch_fill_int(t->B.err + 1, 0, 1, t->stream);
        launch_fill_int(t->B.err + 3, 0, 1, t->stream);
        HIPCHK(hipStreamSynchronize(t->stream));
    }
    return NBX_OK;
}

int nbx_collide(nbx_tree *t, nbx_particles *p, float r_coll, float restitution)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "launchCollisionKernel before summarize");
    DiskTable D;
    memset(&D, 0, sizeof(D));
    for (int s = 0; s < t->T.nseg; ++s) D.d[s].r_coll = r_coll;
    launch_collide(t->T, D, p->vx, p->vy, p->vz, restitution, t->B, t->stream);
    // pass 2 (BHTree.cu:676-686): v += dv
    launch_kick_drag(p->vx, p->vy, p->vz, p->ax, p->ay, p->az, t->B.dv, p->x, p->z, 0.f, 0, 1.f,
                     0.f, p->n, nullptr, t->stream);
    return finish(t);
}

int nbx_perturber_force(nbx_particles *p, float px, float py, float pz, float pmass,
                        float softening)
{
    if (!particles_ok(p)) return fail(NBX_E_ARG, "invalid particle arrays");
    launch_perturber(p->ax, p->ay, p->az, p->x, p->y, p->z, px, py, pz, pmass, softening, p->n,
                     nullptr);
    return finish_plain();
}

int nbx_half_kick(nbx_particles *p, float dt_half)
{
    if (!particles_ok(p)) return fail(NBX_E_ARG, "invalid particle arrays");
    launch_half_kick(p->vx, p->vy, p->vz, p->ax, p->ay, p->az, dt_half, p->n, nullptr);
    return finish_plain();
}

int nbx_drift(nbx_particles *p, float dt)
{
    if (!particles_ok(p)) return fail(NBX_E_ARG, "invalid particle arrays");
    launch_drift(p->x, p->y, p->z, p->vx, p->vy, p->vz, dt, p->n, nullptr);
    return finish_plain();
}

int nbx_reset_acceleration(nbx_particles *p)
{
    if (!particles_ok(p)) return fail(NBX_E_ARG, "invalid particle arrays");
    launch_reset_acc(p->ax, p->ay, p->az, p->n, nullptr);
    return finish_plain();
}

int nbx_gas_drag(nbx_particles *p, float drag_rate, float dt, float star_mass)
{
    if (!particles_ok(p)) return fail(NBX_E_ARG, "invalid particle arrays");
    if (drag_rate <= 0.0f) return NBX_OK;  // Integrator.cu:102
    const float decay = expf(-drag_rate * dt);
    launch_gas_drag(p->x, p->z, p->vx, p->vy, p->vz, decay, star_mass, p->n, nullptr);
    return finish_plain();
}

int nbx_fof(nbx_tree *t, const nbx_particles *p, float link_length, int32_t *dev_labels)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    std::vector<float> ll(t->T.nseg, link_length);
    return nbx_fof_disks(t, p, ll.data(), dev_labels);
}

int nbx_fof_disks(nbx_tree *t, const nbx_particles *p, const float *link_lengths,
                  int32_t *dev_labels)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (!dev_labels || !link_lengths) return fail(NBX_E_ARG, "null labels or link lengths");
    LinkTable LT{};
    for (int s = 0; s < t->T.nseg; ++s) LT.ll[s] = link_lengths[s];
    launch_bbox(t->T, p->x, p->y, p->z, t->B, t->stream);
    launch_root(t->T, t->B, t->stream);
    // the bucket splitters of the tree's last build serve any positions (a
    // bucket they misfit is sorted in global memory), so FoF after steps
    // does not sample new ones
    launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream, !t->B.split_ready);
    launch_summarize(t->T, t->B, t->stream);
    t->stage = kSummarized;
    launch_fof(t->T, LT, dev_labels, t->B, t->stream);
    return finish(t);
}

int nbx_init_forces(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                    const nbx_disk_params *dp)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (!sp || !dp) return fail(NBX_E_ARG, "null parameters");
    if (t->precision == NBX_PRECISION_FP64) {
        rc = fp64_checks(t, sp);
        if (rc) return rc;
        // the FP64 state starts from the caller's float state
        launch_load64(p->x, p->y, p->z, p->vx, p->vy, p->vz, t->B.s64, p->n, t->stream);
        t->s64_ready = true;
        launch_bbox(t->T, p->x, p->y, p->z, t->B, t->stream);
        launch_root(t->T, t->B, t->stream);
        launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream, true, LptDev{},
                     t->B.s64);
        launch_summarize(t->T, t->B, t->stream);
        t->stage = kSummarized;
        rc = upload_steps(t, dp, 0, 1, sp->dt, true);
        if (rc) return rc;
        launch_rec64(t->T, t->B, t->stream);
        launch_walk64(t->T, t->B.dtab, t->B.step_ctr, (double)sp->theta * (double)sp->theta,
                      (double)sp->softening * (double)sp->softening, sp->star_mass, p->ax, p->ay,
                      p->az, t->B, t->stream);
        return finish(t);
    }
    launch_bbox(t->T, p->x, p->y, p->z, t->B, t->stream);
    launch_root(t->T, t->B, t->stream);
    launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream);
    launch_summarize(t->T, t->B, t->stream);
    t->stage = kSummarized;
    WalkArgs W{sp->theta * sp->theta, sp->softening * sp->softening, sp->softening,
               sp->star_mass, t->B.quad, t->walk_tg, p->ax, p->ay, p->az, nullptr, nullptr, nullptr};
    rc = upload_steps(t, dp, 0, 1, sp->dt, true);
    if (rc) return rc;
    launch_walk(t->T, t->B.dtab, t->B.step_ctr, W, t->B, t->stream);
    return finish(t);
}

// The FP64 step (no collisions, monopole walk): KDK, build and walk on the
// FP64 state; the phases are timed like the float step's (no collision pass).
static void enqueue_step64(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp, int e0)
{
    const double dt = sp->dt, dth = 0.5 * dt;
    const int has_drag = sp->gas_drag_rate > 0.0f;
    const double decay = has_drag ? exp(-(double)sp->gas_drag_rate * dt) : 1.0;
    if (e0 >= 0) ev_record(t, e0 + 0);
    launch_kick_drift_bbox_d(t->T, p->x, p->y, p->z, p->vx, p->vy, p->vz, dth, dt, t->B,
                             t->stream);
    if (e0 >= 0) ev_record(t, e0 + 1);
    launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream, /*resample=*/false, LptDev{},
                 t->B.s64);
    launch_summarize(t->T, t->B, t->stream);
    launch_rec64(t->T, t->B, t->stream);
    if (e0 >= 0) ev_record(t, e0 + 2);
    launch_walk64(t->T, t->B.dtab, t->B.step_ctr, (double)sp->theta * (double)sp->theta,
                  (double)sp->softening * (double)sp->softening, sp->star_mass, p->ax, p->ay,
                  p->az, t->B, t->stream);
    if (e0 >= 0) ev_record(t, e0 + 3);
    if (e0 >= 0) ev_record(t, e0 + 4);
    launch_kick_drag_d(p->vx, p->vy, p->vz, dth, has_drag, decay, sp->star_mass, p->n,
                       t->B.step_ctr, t->B, t->stream);
    if (e0 >= 0) ev_record(t, e0 + 5);
}

// A float step whose end can carry the next step's first half-kick and
// drift (k_drag_kick_drift): not with phase timing (it times the kicks
// apart), not with the longest-first wave order (whose counting pass rides
// on the kick-drift kernel).
static bool fusable(const nbx_tree *t)
{
    return t->precision == NBX_PRECISION_FLOAT && !t->lpt && !t->timing;
}

// First half-kick and drift with the boxes of the new positions (main.cu:104-107).
static void enqueue_head(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                         const LptDev &lpt)
{
    const float dt = sp->dt, dth = sp->dt * 0.5f;
    if (t->precision == NBX_PRECISION_COMPENSATED)
        launch_kick_drift_bbox_c(t->T, p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay,
                                 p->az, dth, dt, t->B, t->stream, lpt);
    else
        launch_kick_drift_bbox(t->T, p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay,
                               p->az, dth, dt, t->B, t->stream, lpt);
}

// Rebuild, force and collision pass (main.cu:109-130).
static void enqueue_body(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                         const DiskTable &D, int e0, const LptDev &lpt)
{
    WalkArgs W{sp->theta * sp->theta, sp->softening * sp->softening, sp->softening,
               sp->star_mass, t->B.quad, t->walk_tg, p->ax, p->ay, p->az, nullptr, nullptr, nullptr,
               t->lpt ? 1 : 0, t->wave_order == NBX_WAVES_XCD ? 1 : 0};
    const bool overlap = !t->timing;
    launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream, /*resample=*/false, lpt);
    launch_summarize(t->T, t->B, t->stream);
    if (e0 >= 0) ev_record(t, e0 + 2);
    // The collision pass reads the tree and the velocities and writes only
    // B.dv; the walk writes only the accelerations.  Without phase timing
    // both run as one launch (collision blocks first, then the walk blocks),
    // so the step has no second stream and no cross-stream hand-off.
    if (sp->collisions && overlap) {
        launch_walk_collide(t->T, t->B.dtab, t->B.step_ctr, W, D, p->vx, p->vy, p->vz,
                            sp->restitution, t->B, t->stream);
        if (e0 >= 0) ev_record(t, e0 + 3);
    } else {
        launch_walk(t->T, t->B.dtab, t->B.step_ctr, W, t->B, t->stream);
        if (e0 >= 0) ev_record(t, e0 + 3);
        if (sp->collisions)
            launch_collide(t->T, D, p->vx, p->vy, p->vz, sp->restitution, t->B, t->stream);
    }
}

// Second half-kick with the collision deltas and the drag (main.cu:131-139);
// with `next`, also the next step's first half-kick and drift (fusable trees).
static void enqueue_tail(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp, bool next)
{
    const float dt = sp->dt, dth = sp->dt * 0.5f;
    const int has_drag = sp->gas_drag_rate > 0.0f;
    const float decay = has_drag ? expf(-sp->gas_drag_rate * sp->dt) : 1.0f;  // Integrator.cu:103
    const float *dv = sp->collisions ? t->B.dv : nullptr;
    if (next)
        launch_drag_kick_drift(t->T, p->x, p->y, p->z, p->vx, p->vy, p->vz, p->ax, p->ay, p->az,
                               dv, dth, dt, has_drag, decay, sp->star_mass, t->B.step_ctr, t->B,
                               t->stream);
    else if (t->precision == NBX_PRECISION_COMPENSATED)
        launch_kick_drag_c(p->vx, p->vy, p->vz, p->ax, p->ay, p->az, dv, p->x, p->z, dth,
                           has_drag,
                           has_drag ? exp(-(double)sp->gas_drag_rate * (double)sp->dt) : 1.0,
                           sp->star_mass, p->n, t->B.step_ctr, t->B, t->stream);
    else
        launch_kick_drag(p->vx, p->vy, p->vz, p->ax, p->ay, p->az, dv, p->x, p->z, dth, has_drag,
                         decay, sp->star_mass, p->n, t->B.step_ctr, t->stream);
}

// One whole step (timed phases when e0 >= 0).
static void enqueue_step(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                         const DiskTable &D, int e0)
{
    if (t->precision == NBX_PRECISION_FP64) {
        enqueue_step64(t, p, sp, e0);
        return;
    }
    // the next walk's wave order, from the last walk's wave costs: counted by
    // the kick-drift pass, placed by spare blocks of the key pass
    const LptDev lpt = t->lpt ? lpt_dev(t->B, walk_waves(t->T, t->walk_tg), t->B.step_ctr)
                              : LptDev{};
    if (e0 >= 0) ev_record(t, e0 + 0);
    enqueue_head(t, p, sp, lpt);
    if (e0 >= 0) ev_record(t, e0 + 1);
    enqueue_body(t, p, sp, D, e0, lpt);
    if (e0 >= 0) ev_record(t, e0 + 4);
    enqueue_tail(t, p, sp, false);
    if (e0 >= 0) ev_record(t, e0 + 5);
}

// n steps of a fusable tree: the head once, then every step's body and
// tail, each tail but the last carrying the next step's head.
static void enqueue_steps_fused(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                                const DiskTable &D, int n_steps)
{
    enqueue_head(t, p, sp, LptDev{});
    for (int k = 0; k < n_steps; ++k) {
        enqueue_body(t, p, sp, D, -1, LptDev{});
        enqueue_tail(t, p, sp, k + 1 < n_steps);
    }
}

// Everything a captured step bakes into its kernel arguments.
static std::vector<unsigned char> step_key(const nbx_particles *p, const nbx_step_params *sp,
                                           const DiskTable &D, const Bufs &B, int targets)
{
    std::vector<unsigned char> k(sizeof(*p) + sizeof(*sp) + sizeof(D) + sizeof(void *) + 2);
    unsigned char *q = k.data();
    memcpy(q, p, sizeof(*p));
    q += sizeof(*p);
    memcpy(q, sp, sizeof(*sp));
    q += sizeof(*sp);
    memcpy(q, &D, sizeof(D));
    q += sizeof(D);
    memcpy(q, &B.dtab, sizeof(void *));
    q += sizeof(void *);
    q[0] = B.lo != nullptr;
    q[1] = (unsigned char)targets;
    return k;
}

// Per-step disk table of the next n_steps steps and the walk's wave-order
// buffers, on the tree's own stream.
static int prepare_steps(nbx_tree *t, const nbx_step_params *sp, const nbx_disk_params *dp,
                         int32_t step0, int32_t n_steps, DiskTable &D)
{
    if (t->precision == NBX_PRECISION_FP64) {
        int rc = fp64_checks(t, sp);
        if (rc) return rc;
        if (!t->s64_ready)
            return fail(NBX_E_STATE, "FP64 state not loaded: call nbx_init_forces after "
                                     "nbx_tree_set_precision(t, NBX_PRECISION_FP64)");
    }
    int rc = upload_steps(t, dp, step0, n_steps, sp->dt, false);
    if (rc) return rc;
    D = disk_rcoll(t, dp);
    if (t->lpt && t->lpt_tg != t->walk_tg) {  // wave orders of another shape: start over
        reset_lpt(t->B, t->stream);
        t->lpt_tg = t->walk_tg;
    }
    return NBX_OK;
}

static int capture(nbx_tree *t, hipGraphExec_t *out, const std::function<void()> &enqueue)
{
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    HIPCHK(hipStreamEndCapture(t->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(NBX_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return NBX_OK;
}

// The tree's captured step (kept while nothing it bakes in changes): one
// graph of the whole step, or for a fusable tree the step whose end carries
// the next step's head (`graph`) and the last step of a call (`graph_last`).
static int ensure_step_graph(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
                             const DiskTable &D)
{
    const bool fuse = fusable(t);
    std::vector<unsigned char> key = step_key(p, sp, D, t->B, t->walk_tg);
    key.push_back(fuse ? 1 : 0);
    if (t->graph && key == t->graph_key) return NBX_OK;
    HIPCHK(drop_graphs(t));
    int rc;
    if (fuse) {
        rc = capture(t, &t->graph, [&] {
            enqueue_body(t, p, sp, D, -1, LptDev{});
            enqueue_tail(t, p, sp, true);
        });
        if (!rc)
            rc = capture(t, &t->graph_last, [&] {
                enqueue_body(t, p, sp, D, -1, LptDev{});
                enqueue_tail(t, p, sp, false);
            });
    } else {
        rc = capture(t, &t->graph, [&] { enqueue_step(t, p, sp, D, -1); });
    }
    if (rc) {
        (void)drop_graphs(t);
        return rc;
    }
    t->fused = fuse;
    t->graph_key = key;
    return NBX_OK;
}

// Step k of n_steps of a tree whose graphs are ready (the head before step
// 0 when the steps are fused).
static int launch_graph_step(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp, int k,
                             int n_steps)
{
    if (!t->fused) {
        HIPCHK(hipGraphLaunch(t->graph, t->stream));
        return NBX_OK;
    }
    if (k == 0) enqueue_head(t, p, sp, LptDev{});
    HIPCHK(hipGraphLaunch(k + 1 < n_steps ? t->graph : t->graph_last, t->stream));
    return NBX_OK;
}

int nbx_step(nbx_tree *t, nbx_particles *p, const nbx_step_params *sp,
             const nbx_disk_params *dp, int32_t step0, int32_t n_steps)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (!sp || !dp) return fail(NBX_E_ARG, "null parameters");
    if (n_steps < 0 || step0 < 0) return fail(NBX_E_ARG, "negative step range");
    if (n_steps == 0) return finish(t);
    DiskTable D;
    rc = prepare_steps(t, sp, dp, step0, n_steps, D);
    if (rc) return rc;
    const bool graph = sp->use_graph && !t->timing;
    if (graph) {
        rc = ensure_step_graph(t, p, sp, D);
        if (rc) return rc;
        for (int k = 0; k < n_steps; ++k)
            if ((rc = launch_graph_step(t, p, sp, k, n_steps))) return rc;
    } else if (fusable(t)) {
        enqueue_steps_fused(t, p, sp, D, n_steps);
    } else {
        const int nev = 6 * n_steps;
        if (t->timing && (int)t->ev.size() < nev) {
            const size_t old = t->ev.size();
            t->ev.resize(nev);
            for (size_t q = old; q < t->ev.size(); ++q) HIPCHK(hipEventCreate(&t->ev[q]));
        }
        for (int k = 0; k < n_steps; ++k) enqueue_step(t, p, sp, D, t->timing ? 6 * k : -1);
        if (t->timing) {
            float acc_ms[5] = {0, 0, 0, 0, 0};
            HIPCHK(hipEventSynchronize(t->ev[nev - 1]));
            for (int k = 0; k < n_steps; ++k)
                for (int q = 0; q < 5; ++q) {
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, t->ev[6 * k + q], t->ev[6 * k + q + 1]));
                    acc_ms[q] += ms;
                }
            for (int q = 0; q < 5; ++q) t->phase_ms[q] = acc_ms[q] / n_steps;
            t->phase_n = n_steps;
        }
    }
    t->stage = kSummarized;
    return finish(t);
}

// Several trees (each one disk or a batch of disks) stepped together: each
// tree replays its own captured step (the graph nbx_step captures for it) on
// its own stream, step by step, so the GPU runs one tree's latency-bound
// build passes beside another tree's walk.  The trees' streams start after
// the work already on trees[0]'s stream, and trees[0]'s stream waits for all
// of them.  Every tree computes exactly what nbx_step computes for it alone.
int nbx_step_concurrent(nbx_tree *const *trees, nbx_particles *const *ps, int32_t n_trees,
                        const nbx_step_params *sp, const nbx_disk_params *const *dps,
                        int32_t step0, int32_t n_steps)
{
    if (!trees || !ps || !dps || !sp) return fail(NBX_E_ARG, "null arguments");
    if (n_trees < 1 || n_trees > NBX_MAX_GROUP)
        return fail(NBX_E_ARG, "tree count %d not in [1,%d]", n_trees, NBX_MAX_GROUP);
    if (n_steps < 0 || step0 < 0) return fail(NBX_E_ARG, "negative step range");
    for (int i = 0; i < n_trees; ++i) {
        int rc = check_tree_particles(trees[i], ps[i]);
        if (rc) return rc;
        if (!dps[i]) return fail(NBX_E_ARG, "null disk parameters of tree %d", i);
        for (int j = 0; j < i; ++j) {
            if (trees[j] == trees[i]) return fail(NBX_E_ARG, "tree %d appears twice", i);
            // the trees run side by side: two of them on one state would race
            if (ps[j]->x == ps[i]->x || ps[j]->vx == ps[i]->vx || ps[j]->ax == ps[i]->ax)
                return fail(NBX_E_ARG, "trees %d and %d step the same particle arrays", j, i);
        }
    }
    nbx_tree *lead = trees[0];
    HIPCHK(hipEventRecord(lead->gord, lead->stream));
    for (int i = 1; i < n_trees; ++i) HIPCHK(hipStreamWaitEvent(trees[i]->stream, lead->gord, 0));
    bool direct = !sp->use_graph;
    for (int i = 0; i < n_trees; ++i) direct = direct || trees[i]->timing;
    if (direct || n_steps == 0) {  // no graphs: each tree steps on its own stream in turn
        for (int i = 0; i < n_trees; ++i) {
            int rc = nbx_step(trees[i], ps[i], sp, dps[i], step0, n_steps);
            if (rc) return rc;
        }
        for (int i = 1; i < n_trees; ++i) {
            HIPCHK(hipEventRecord(trees[i]->gord, trees[i]->stream));
            HIPCHK(hipStreamWaitEvent(lead->stream, trees[i]->gord, 0));
        }
        return NBX_OK;
    }
    for (int i = 0; i < n_trees; ++i) {
        nbx_tree *t = trees[i];
        DiskTable D;
        int rc = prepare_steps(t, sp, dps[i], step0, n_steps, D);
        if (rc) return rc;
        rc = ensure_step_graph(t, ps[i], sp, D);
        if (rc) return rc;
    }
    for (int k = 0; k < n_steps; ++k)
        for (int i = 0; i < n_trees; ++i) {
            const int rc = launch_graph_step(trees[i], ps[i], sp, k, n_steps);
            if (rc) return rc;
        }
    for (int i = 1; i < n_trees; ++i) {
        HIPCHK(hipEventRecord(trees[i]->gord, trees[i]->stream));
        HIPCHK(hipStreamWaitEvent(lead->stream, trees[i]->gord, 0));
    }
    for (int i = 0; i < n_trees; ++i) {
        trees[i]->stage = kSummarized;
        int rc = finish(trees[i]);
        if (rc) return rc;
    }
    return NBX_OK;
}

// ---------------------------------------------------------------------------
// Reverse mode of the collision-free step (differentiable window)
// ---------------------------------------------------------------------------

int nbx_tree_set_adjoint(nbx_tree *t, int enable)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (!enable || t->adjoint) {
        t->adjoint = t->adjoint && enable;
        return NBX_OK;
    }
    Bufs &B = t->B;
    const int n = t->T.n_total;
    int rc;
    if ((rc = dmalloc(&B.rec_r, B.n_rec)) ||
        (rc = dmalloc(&B.adj_g, 3 * (size_t)B.n_rec)) ||
        (rc = dmalloc(&B.adj_d, 3 * (size_t)(n + 1))) ||
        (rc = dmalloc(&B.adj_blk, 3 * (size_t)adj_scan_blocks(n))) || (rc = dmalloc(&B.adj_p, 4)))
        return rc;
    HIPCHK(hipMemset(B.rec_r, 0, sizeof(int2) * B.n_rec));
    HIPCHK(hipHostMalloc((void **)&t->h_adj, 4 * sizeof(double)));
    HIPCHK(drop_graphs(t));  // the captured step holds the build without range outputs
    t->adjoint = true;
    return NBX_OK;
}

}  // extern "C"

namespace {

int adj_check(nbx_tree *t, const nbx_particles *p)
{
    int rc = check_tree_particles(t, p);
    if (rc) return rc;
    if (!t->adjoint) return fail(NBX_E_STATE, "reverse mode not enabled (nbx_tree_set_adjoint)");
    if (t->T.nseg != 1) return fail(NBX_E_ARG, "reverse mode takes one disk per tree");
    if (t->B.quad) return fail(NBX_E_ARG, "reverse mode differentiates the monopole walk only");
    if (t->precision != NBX_PRECISION_FLOAT)
        return fail(NBX_E_ARG, "reverse mode follows the float KDK only");
    return NBX_OK;
}

// Perturber of the force evaluation at sim_time (Perturber.h:60-66 in float,
// as the forward step) and dX_p/dr_p of the orbit r (cos phi, 0, sin phi),
// phi = phase0 + omega t, omega = sqrt(M / r^3) (Perturber.h:39-58).
AdjWalkArgs adj_args(const nbx_tree *t, float theta, float softening, float star_mass,
                     const nbx_disk_params *dp, float sim_time, const double *ab, double *xb)
{
    AdjWalkArgs A{};
    A.theta2 = theta * theta;
    A.eps2 = softening * softening;
    A.softening = softening;
    A.star_mass = star_mass;
    A.targets = t->walk_tg;
    A.ab = ab;
    A.xb = xb;
    if (dp && dp->perturber_enabled) {
        const float phi = dp->pert_phase0 + dp->pert_omega * sim_time;
        A.px = dp->pert_radius_au * cosf(phi);
        A.py = 0.f;
        A.pz = dp->pert_radius_au * sinf(phi);
        A.pmass = dp->pert_mass_msun;
        const double r = dp->pert_radius_au, w = dp->pert_omega, ph = phi;
        const double dphi = (double)sim_time * (-1.5 * w / r);
        A.drx = (float)(cos(ph) - r * sin(ph) * dphi);
        A.dry = 0.f;
        A.drz = (float)(sin(ph) + r * cos(ph) * dphi);
    }
    return A;
}

// Adjoint force over the tree of p's positions; pgrad += dL/d(m_p, r_p).
int adj_force(nbx_tree *t, const nbx_particles *p, const AdjWalkArgs &A, double *pgrad)
{
    launch_bbox(t->T, p->x, p->y, p->z, t->B, t->stream);
    launch_root(t->T, t->B, t->stream);
    launch_build(t->T, p->x, p->y, p->z, p->mass, t->B, t->stream, !t->B.split_ready);
    launch_summarize(t->T, t->B, t->stream);
    t->stage = kSummarized;
    launch_adj_force(t->T, A, t->B, t->stream);
    HIPCHK(hipMemcpyAsync(t->h_adj, t->B.adj_p, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          t->stream));
    const bool was_async = t->async;
    t->async = false;
    const int rc = finish(t);
    t->async = was_async;
    if (rc) return rc;
    if (pgrad) {
        pgrad[0] += t->h_adj[0];
        pgrad[1] += t->h_adj[1];
    }
    return NBX_OK;
}

}  // namespace

extern "C" {

int nbx_adj_force(nbx_tree *t, const nbx_particles *p, float theta, float softening,
                  float star_mass, const nbx_disk_params *dp, float sim_time, const double *abar,
                  double *xbar, double *pgrad)
{
    int rc = adj_check(t, p);
    if (rc) return rc;
    if (!abar || !xbar) return fail(NBX_E_ARG, "null adjoint arrays");
    return adj_force(t, p, adj_args(t, theta, softening, star_mass, dp, sim_time, abar, xbar),
                     pgrad);
}

int nbx_adj_init_forces(nbx_tree *t, const nbx_particles *s0, const nbx_step_params *sp,
                        const nbx_disk_params *dp, const double *abar, double *xbar,
                        double *pgrad)
{
    int rc = adj_check(t, s0);
    if (rc) return rc;
    if (!sp || !dp) return fail(NBX_E_ARG, "null parameters");
    if (!abar || !xbar) return fail(NBX_E_ARG, "null adjoint arrays");
    // initial forces use the perturber at t = 0 (main.cu:85-96)
    return adj_force(t, s0, adj_args(t, sp->theta, sp->softening, sp->star_mass, dp, 0.0f, abar,
                                     xbar),
                     pgrad);
}

int nbx_adj_step(nbx_tree *t, const nbx_particles *s0, const nbx_particles *s1,
                 const nbx_step_params *sp, const nbx_disk_params *dp, int32_t step,
                 double *xbar, double *vbar, double *abar, double *pgrad)
{
    int rc = adj_check(t, s1);
    if (rc) return rc;
    if (!particles_ok(s0) || s0->n != s1->n) return fail(NBX_E_ARG, "invalid step-start state");
    if (!sp || !dp) return fail(NBX_E_ARG, "null parameters");
    if (!xbar || !vbar || !abar) return fail(NBX_E_ARG, "null adjoint arrays");
    if (sp->collisions) return fail(NBX_E_ARG, "reverse mode needs collisions off");
    if (step < 0) return fail(NBX_E_ARG, "negative step index");
    const int n = s1->n;
    const float dt = sp->dt, dth = sp->dt * 0.5f;
    const int has_drag = sp->gas_drag_rate > 0.0f;
    const float decay = has_drag ? expf(-sp->gas_drag_rate * sp->dt) : 1.0f;  // Integrator.cu:103
    launch_adj_kick_drag(s1->x, s1->z, s0->vx, s0->vy, s0->vz, s0->ax, s0->ay, s0->az, s1->ax,
                         s1->ay, s1->az, dth, has_drag, decay, sp->star_mass, xbar, vbar, abar,
                         n, t->stream);
    const float sim_time = (float)(step + 1) * dt;  // main.cu:110
    rc = adj_force(t, s1, adj_args(t, sp->theta, sp->softening, sp->star_mass, dp, sim_time,
                                   abar, xbar),
                   pgrad);
    if (rc) return rc;
    launch_adj_drift_kick(xbar, vbar, abar, dt, dth, n, t->stream);
    const bool was_async = t->async;
    t->async = false;
    rc = finish(t);
    t->async = was_async;
    return rc;
}

int nbx_tree_stats(nbx_tree *t, int32_t *n_nodes, int32_t *max_level)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    if (t->stage < kSummarized) return fail(NBX_E_STATE, "statistics before summarize");
    std::vector<int> alloc(kMaxSeg), lev(kMaxSeg);
    launch_tree_stats(t->T, t->B, t->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->stream));
    HIPCHK(hipMemcpy(alloc.data(), t->B.alloc, sizeof(int) * kMaxSeg, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lev.data(), t->B.maxlev, sizeof(int) * kMaxSeg, hipMemcpyDeviceToHost));
    for (int s = 0; s < t->T.nseg; ++s) {
        if (n_nodes) n_nodes[s] = alloc[s] + 1;  // child records + the root
        if (max_level) max_level[s] = lev[s];
    }
    return NBX_OK;
}

int nbx_tree_set_timing(nbx_tree *t, int on)
{
    if (!t) return fail(NBX_E_ARG, "null tree");
    t->timing = on != 0;
    return NBX_OK;
}

int nbx_probe_fp32(int kind, double *tflops)
{
    if (!tflops || (kind != 0 && kind != 1)) return fail(NBX_E_ARG, "probe kind %d", kind);
    float *scratch = nullptr;
    HIPCHK(hipMalloc((void **)&scratch, 256 * 4 * 8 * sizeof(float)));
    const double v = probe_fp32(kind, scratch, nullptr);
    HIPCHK(hipDeviceSynchronize());
    (void)hipFree(scratch);
    if (v < 0.0) return fail(NBX_E_HIP, "probe timing failed");
    *tflops = v;
    return NBX_OK;
}

int nbx_phase_ms(nbx_tree *t, const char *phase, float *ms)
{
    if (!t || !phase || !ms) return fail(NBX_E_ARG, "null argument");
    for (int q = 0; q < 5; ++q)
        if (!strcmp(phase, kPhases[q])) {
            *ms = t->phase_ms[q];
            return NBX_OK;
        }
    return fail(NBX_E_ARG, "unknown phase '%s'", phase);
}

}  // extern "C"

// ===========================================================================
// FoF labels without allocation (static device scratch)
// ===========================================================================
namespace {

// Bytes of the static scratch: the core layout of one disk of
// NBX_NOALLOC_MAX_BODIES bodies (about 570 bytes per body, 384 of them the
// walk records' 8-record blocks, plus the fixed histogram and padding) fits
// with margin; checked when a device first uses it.
constexpr size_t kNoAllocScratchBytes = (size_t)40 << 20;
constexpr int kMaxDevices = 64;

}  // namespace

// One copy per GPU, alive until the process ends.  Host code never uses the
// name itself (a host-side stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Bump layout of the core buffers over `base` (nullptr: sizes only, every
// pointer stays null); returns the bytes used.  Each buffer starts on a
// 256-byte boundary.
size_t carve_core(const SegTable &T, Bufs &B, unsigned char *base)
{
    size_t off = 0;
    auto take = [&](auto *&ptr, size_t count) {
        using Elem = std::remove_reference_t<decltype(*ptr)>;
        off = (off + 255) & ~(size_t)255;
        ptr = base ? reinterpret_cast<Elem *>(base + off) : nullptr;
        off += (count ? count : 1) * sizeof(Elem);
    };
    core_layout(T, B, take);
    if (base) scan_views(B);
    return off;
}

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_of_current_device(unsigned char **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(NBX_E_ARG, "device %d not supported", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(NBX_E_HIP, "static scratch has no device address");
        // the largest disk the entry point accepts must fit
        SegTable T;
        const int32_t nmax = NBX_NOALLOC_MAX_BODIES;
        int rc = make_segtable(1, &nmax, T);
        if (rc) return rc;
        Bufs B;
        const size_t need = carve_core(T, B, nullptr);
        if (need > kNoAllocScratchBytes)
            return fail(NBX_E_STATE, "static scratch of %zu bytes holds %zu needed (internal)",
                        kNoAllocScratchBytes, need);
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return NBX_OK;
}

}  // namespace

extern "C" int nbx_fof_noalloc(const nbx_particles *p, float link_length, int32_t *dev_labels,
                               void *stream)
{
    if (!particles_ok(p)) return fail(NBX_E_ARG, "nbx_fof_noalloc: invalid particle arrays");
    if (!dev_labels) return fail(NBX_E_ARG, "nbx_fof_noalloc: null labels");
    if (p->n > NBX_NOALLOC_MAX_BODIES)
        return fail(NBX_E_ARG,
                    "nbx_fof_noalloc: %d bodies exceed the static scratch's maximum of %d "
                    "(use nbx_fof with a tree)",
                    p->n, NBX_NOALLOC_MAX_BODIES);
    SegTable T;
    int rc = make_segtable(1, &p->n, T);
    if (rc) return rc;
    unsigned char *base = nullptr;
    if ((rc = scratch_of_current_device(&base)) != NBX_OK) return rc;
    Bufs B;
    if (carve_core(T, B, base) > kNoAllocScratchBytes)
        return fail(NBX_E_STATE, "nbx_fof_noalloc: scratch layout overflow (internal)");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the previous call's contents stay in the scratch: the error flags
    // start from zero (the key pass resets its own tickets)
    launch_fill_int(B.err, 0, 4, st);
    launch_bbox(T, p->x, p->y, p->z, B, st);
    launch_root(T, B, st);
    launch_build(T, p->x, p->y, p->z, p->mass, B, st);
    launch_summarize(T, B, st);
    LinkTable LT{};
    LT.ll[0] = link_length;
    launch_fof(T, LT, dev_labels, B, st);
    HIPCHK(hipGetLastError());
    int h_err[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(h_err, B.err, sizeof(h_err), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h_err[0] & kErrTree) return fail(NBX_E_STATE, "inconsistent octree (internal error)");
    return NBX_OK;
}
