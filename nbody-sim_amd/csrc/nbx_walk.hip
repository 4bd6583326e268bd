// nbx_walk.hip -- tree walks on gfx950: gravity, collision impulse, FoF.
//
// Gravity (wave-coherent, nbx_wave.hpp): a wavefront owns TG consecutive
// targets in Morton order and walks the compressed octree with a
// wave-uniform stack in LDS; a target evaluates a record only if its own
// walk reached it, so every target visits exactly the nodes of its own
// per-body walk (BHTree.cu:555-593).
//
// Collision and FoF (per lane): the search radius is far below the particle
// spacing, so the per-body paths diverge immediately; each lane climbs to
// the lowest cell that holds its search sphere and descends below it
// without a stack (octree parent links), pruning cells by the exact
// distance to the cell box.
#include "nbx_wave.hpp"

namespace nbx {

// Everything the gravity walk reads besides the disk tables.
struct GravArgs {
    const DiskDev *dtab;        // per-step perturber of every disk (or null)
    const int *step_ctr;
    const float4 *ra;
    const int4 *rb;
    const float4 *rq;
    const float4 *bodies;
    const int *perm;
    float theta2, eps2, softening, star_mass;
    float *ax, *ay, *az, *sax, *say, *saz;
    int *err;
    const int *order;           // longest-first wave order (or null)
    int *cost;
    int lpt_stride;
    int xcd_chunk;              // > 0: contiguous wave ranges per XCD
};

template <int TG, bool QUAD>
using GravLds = std::conditional_t<QUAD, WaveLdsQT<(TG > 32)>, WaveLdsT<(TG > 32)>>;

// The four waves of walk block `block`.
template <int TG, bool QUAD, int U, bool SELF>
__device__ inline void walk_block(const SegTable &T, const WaveMap &M, const GravArgs &A,
                                  int block, GravLds<TG, QUAD> *lds)
{
    const DiskDev *__restrict__ dtab = A.dtab;
    const int *__restrict__ step_ctr = A.step_ctr;
    const float4 *__restrict__ ra = A.ra;
    const int4 *__restrict__ rb = A.rb;
    const float4 *__restrict__ rq = A.rq;
    const float4 *__restrict__ bodies = A.bodies;
    const int *__restrict__ perm = A.perm;
    const float theta2 = A.theta2, eps2 = A.eps2, softening = A.softening,
                star_mass = A.star_mass;
    float *ax = A.ax, *ay = A.ay, *az = A.az, *sax = A.sax, *say = A.say, *saz = A.saz;
    int *err = A.err;
    const int *__restrict__ order = A.order;
    int *__restrict__ cost = A.cost;
    const int lpt_stride = A.lpt_stride, xcd_chunk = A.xcd_chunk;
    const int wib = threadIdx.x >> 6;
    const int slot = block * kGravWaves + wib;
    // launch slot -> target group: longest waves of the previous step first,
    // or (xcd_chunk > 0) contiguous Morton ranges per XCD: workgroups are
    // dispatched to the 8 XCDs round robin, so block % 8 is the XCD and its
    // waves share the tree records of one region in that XCD's L2
    int w;
    if (xcd_chunk > 0) {
        const int j = (block / kXcds) * kGravWaves + wib;
        if (j >= xcd_chunk) return;
        w = (block % kXcds) * xcd_chunk + j;
        if (w >= M.wave[T.nseg]) return;  // uniform per wave
    } else {
        if (slot >= M.wave[T.nseg]) return;
        w = order ? order[(*step_ctr & 1) * lpt_stride + slot] : slot;
    }
    const int par = order ? (*step_ctr & 1) * lpt_stride : 0;
    const int s = seg_of_wmap(T, M, w);
    const int lane = __lane_id(), tt = lane % TG, rr = lane / TG;
    const int pos = T.off[s] + (w - M.wave[s]) * TG + tt;
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    GravityVisitor<QUAD, SELF> v{b.x, b.y, b.z, 1.0f / theta2, eps2, pos, 0.f, 0.f, 0.f};
    const uint64_t active = ballot64(valid && rr == 0);  // bit tt per valid target
    const int work = wave_walk<TG, QUAD, U>(ra, rb, rq, T.noff[s], active, lds[wib], err, v);
    if (cost && lane == 0) cost[par + w] = work;
    const float fx = slot_sum<TG>(v.fx), fy = slot_sum<TG>(v.fy), fz = slot_sum<TG>(v.fz);
    if (!valid || rr != 0) return;
    const int i = perm[pos];
    float ox, oy, oz;
    {
#pragma clang fp contract(off)
        // star at the origin (BHTree.cu:547-553)
        float dx = -b.x, dy = -b.y, dz = -b.z;
        float r2 = dx * dx + dy * dy + dz * dz + softening * softening;
        float r = sqrtf(r2);
        float f = star_mass / (r2 * r);
        ox = f * dx + fx;
        oy = f * dy + fy;
        oz = f * dz + fz;
        DiskDev d{};
        if (dtab) d = dtab[*step_ctr * kMaxSeg + s];
        if (d.pmass > 0.f)
            perturber_add_dev(b.x, b.y, b.z, d.px, d.py, d.pz, d.pmass, softening, ox, oy, oz);
    }
    ax[i] = ox;
    ay[i] = oy;
    az[i] = oz;
    if (sax) {
        sax[i] = fx;
        say[i] = fy;
        saz[i] = fz;
    }
}

template <int TG, bool QUAD, int U, bool SELF>
__global__ __launch_bounds__(kGravBlock) void k_walk(SegTable T, WaveMap M, GravArgs A)
{
    __shared__ GravLds<TG, QUAD> lds[kGravWaves];
    walk_block<TG, QUAD, U, SELF>(T, M, A, blockIdx.x, lds);
}

template <int TG>
__global__ __launch_bounds__(kGravBlock) void k_walk_counts(
    SegTable T, WaveMap M, const float4 *__restrict__ ra, const int4 *__restrict__ rb,
    const float4 *__restrict__ bodies, const int *__restrict__ perm, float theta2, int *n_acc,
    int *n_open, int *err, int *wstat)
{
    __shared__ WaveLdsT<(TG > 32)> lds[kGravWaves];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kGravWaves + wib;
    if (w >= M.wave[T.nseg]) return;
    const int s = seg_of_wmap(T, M, w);
    const int lane = __lane_id(), tt = lane % TG, rr = lane / TG;
    const int pos = T.off[s] + (w - M.wave[s]) * TG + tt;
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    CountVisitor v{b.x, b.y, b.z, 1.0f / theta2, pos, 0, 0};
    wave_walk<TG, false>(ra, rb, nullptr, T.noff[s], ballot64(valid && rr == 0), lds[wib], err, v,
                         wstat ? wstat + 3 * w : nullptr);
    const int na = slot_sum_i<TG>(v.n_acc), no = slot_sum_i<TG>(v.n_open);
    if (!valid || rr != 0) return;
    const int i = perm[pos];
    if (n_acc) n_acc[i] = na;
    if (n_open) n_open[i] = no;
}

constexpr int kLStack = 32;   // per-lane stack entries of the FoF searches (LDS)
// Per-lane stack entries of the collision searches, alone or inside the walk
// launch (search radii are tiny: a stack of 12 blocks holds every search
// but those below dense clumps, which go on without it).
constexpr int kLStackC = 12;
// Body waves per collision wave in the fused walk launch: fewer collision
// blocks leave the compute units' slots to the walk's blocks.
constexpr int kCollideRounds = 2;

// Tree arrays of the per-lane searches: octree parents (binary index) of
// octree nodes and of body leaves, every octree node's child block and the
// codes of its children (kids[8 j + k] is the node or ~body of record
// first[j] + k), and the cell boxes of the binary nodes.
struct SearchTree {
    const int *opar, *opar_leaf, *first, *cnt, *kids;
    const float4 *bin_cell;
};
static inline SearchTree search_bufs(const Bufs &B)
{
    return SearchTree{B.opar, B.opar_leaf, B.first, B.cnt, B.kids, B.bin_cell};
}

// ---------------------------------------------------------------------------
// Per-lane descent for small search radii.  A cell record {centre, h} holds
// every body classified into it within h plus the rounding of the float
// descent that produced its centre: at most 21 roundings of half an ulp of
// |root centre| + root half size, bounded by `guard` = 2e-6 of that scale;
// the cell is opened when the target lies within `radius` of that box.
// ---------------------------------------------------------------------------
__device__ inline bool near_cell(float tx, float ty, float tz, const float4 &c, float radius,
                                 float guard)
{
    const float slack = c.w + guard;
    const float ex = fmaxf(fabsf(tx - c.x) - slack, 0.f);
    const float ey = fmaxf(fabsf(ty - c.y) - slack, 0.f);
    const float ez = fmaxf(fabsf(tz - c.z) - slack, 0.f);
    const float r = radius * 1.0001f;
    return ex * ex + ey * ey + ez * ez <= r * r;
}

// Pair(p, j_sorted): called for every leaf within the pruned walk.
__device__ inline float root_guard(const Root &R)
{
    return (fabsf(R.c[0]) + fabsf(R.c[1]) + fabsf(R.c[2]) + R.c[3]) * 2e-6f;
}

// Lowest octree node (binary index) whose cell holds every body within
// `radius` of the target at sorted position pos: the climb from the body's
// own cell stops at the first cell whose box, shrunk by twice the descent
// rounding, contains the search sphere (such a body is classified into that
// cell), or at the root.  -1: the body is its disk's only body.
__device__ inline int search_start(const int *__restrict__ opar, const int *__restrict__ opar_leaf,
                                   const float4 *__restrict__ bin_cell, int pos, float tx,
                                   float ty, float tz, float radius, float guard)
{
    int j = opar_leaf[pos];
    if (j < 0) return -1;
    const float reach = radius * 1.0001f + 2.f * guard;
    for (;;) {
        // one round trip per level: the whole cell and the parent are read
        // together, and the test does not short-circuit
        const float4 c = bin_cell[j];
        const int up = opar[j];
        const bool inside = (fabsf(tx - c.x) + reach < c.w) & (fabsf(ty - c.y) + reach < c.w) &
                            (fabsf(tz - c.z) + reach < c.w);
        if (inside || up < 0) return j;
        j = up;
    }
}

// A record as the per-lane searches read it: the cell {centre | body, h}
// and {first child | sorted position, child count} (rec_b.z / .w, one 64-bit
// load); the record's binary node (rec_b.x) is read only where a search
// needs it, so a chunk of 8 records costs 48 registers.
__device__ inline int2 rec_link(const int4 *__restrict__ rb, int r)
{
    return reinterpret_cast<const int2 *>(rb)[2 * r + 1];
}

// Depth-first search, in pre-order with children in block (Morton) order,
// of the records below octree node `top` (binary index), without a stack:
// the lane keeps the node whose child block it scans and the slot in that
// block; a near internal record is entered at once, and a finished block is
// left through the node's octree parent (opar) at the slot after the node's
// own (found among the parent's child codes).  Any depth is handled.  The
// records of a block are read in chunks of 8, one round trip per chunk.
template <class Pair>
__device__ inline void lane_walk_subtree(const float4 *__restrict__ rc,
                                         const int4 *__restrict__ rb, const SearchTree &S,
                                         int top, float tx, float ty, float tz, float radius,
                                         float guard, Pair &pair)
{
    int node = top, f = S.first[top], c = S.cnt[top], k = 0;
    for (;;) {
        if (k >= c) {  // block done: back to the parent's next slot
            if (node == top) return;
            const int up = S.opar[node];
            const int4 k0 = *reinterpret_cast<const int4 *>(S.kids + 8 * up);
            const int4 k1 = *reinterpret_cast<const int4 *>(S.kids + 8 * up + 4);
            const int fu = S.first[up], cu = S.cnt[up];
            const int at = k0.x == node   ? 0
                           : k0.y == node ? 1
                           : k0.z == node ? 2
                           : k0.w == node ? 3
                           : k1.x == node ? 4
                           : k1.y == node ? 5
                           : k1.z == node ? 6
                                          : 7;
            node = up;
            f = fu;
            c = cu;
            k = at + 1;
            continue;
        }
        int2 L[8];
        float4 Cc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            L[q] = rec_link(rb, f + k + q);  // records past the block exist (padding / later blocks)
            Cc[q] = rc[f + k + q];
        }
        const int kend = min(c, k + 8);
        int down = -1;
        int2 d = make_int2(0, 0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (k + q >= kend || down >= 0) break;
            if (L[q].y == 0) {
                pair(Cc[q], L[q].x);
            } else if (near_cell(tx, ty, tz, Cc[q], radius, guard)) {
                down = q;
                d = L[q];
            }
        }
        if (down < 0) {
            k = kend;
            continue;
        }
        node = rb[f + k + down].x;  // the record's own binary node
        f = d.x;
        c = d.y;
        k = 0;
    }
}

// Depth-first search of the records below one children block {first,
// count} on a per-lane stack of CAP blocks in LDS: a block is scanned
// whole (one round trip per 8 records) and the near internal records among
// its records are pushed.  A lane whose stack is full (a clump with every
// cell in range: up to 7 pending blocks per level) searches those records'
// subtrees without a stack once the chunk is scanned, so no depth or clump
// is out of reach.
template <int CAP, class Pair>
__device__ inline void lane_walk_block(const float4 *__restrict__ rc, const int4 *__restrict__ rb,
                                       const SearchTree &S, int first0, int count0, float tx,
                                       float ty, float tz, float radius, float guard, int2 *st,
                                       int *err, Pair &pair)
{
    // st: this lane's stack, stride kWave between entries (bank-conflict free)
    st[0] = make_int2(first0, count0);
    int sp = 1;
    while (sp > 0) {
        --sp;
        const int2 e = st[sp * kWave];
        const int first = e.x, count = e.y;
        for (int k0 = 0; k0 < count; k0 += 8) {
            unsigned over = 0;  // near internal records that found the stack full
            {
                int2 L[8];
                float4 Cc[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    L[k] = rec_link(rb, first + k0 + k);
                    Cc[k] = rc[first + k0 + k];
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (k0 + k >= count) break;
                    if (L[k].y == 0) {
                        pair(Cc[k], L[k].x);
                    } else if (near_cell(tx, ty, tz, Cc[k], radius, guard)) {
                        if (sp < CAP) {
                            st[sp * kWave] = L[k];
                            ++sp;
                        } else {
                            over |= 1u << k;
                        }
                    }
                }
            }
            while (over) {
                const int k = __builtin_ctz(over);
                over &= over - 1;
                atomicAdd(err + 3, 1);  // diagnostic: searches past the stack
                lane_walk_subtree(rc, rb, S, rb[first + k0 + k].x, tx, ty, tz, radius, guard,
                                  pair);
            }
        }
    }
}

// Collision pass 1 (BHTree.cu:628-672) on the exact pair set
// dist^2 < r_coll^2: J = -(1+e) dvn / 2 for approaching pairs (dvn < 0),
// dv_i -= J n, accumulated in the walk's (Morton) order.
struct CollidePair {
    float bx, by, bz, rc2, e;
    float vix, viy, viz;
    int self;
    const int *perm;
    const float *vx, *vy, *vz;
    float dvx, dvy, dvz;
    int pairs;
    __device__ inline void operator()(const float4 &c, int p)
    {
#pragma clang fp contract(off)
        if (p == self) return;
        const float dx = c.x - bx, dy = c.y - by, dz = c.z - bz;
        const float dist2 = dx * dx + dy * dy + dz * dz;
        if (!(dist2 < rc2)) return;
        ++pairs;
        const int j = perm[p];
        const float dist = sqrtf(dist2) + 1e-10f;
        const float nx = dx / dist, ny = dy / dist, nz = dz / dist;
        const float dvn = (vx[j] - vix) * nx + (vy[j] - viy) * ny + (vz[j] - viz) * nz;
        if (dvn < 0.f) {
            const float J = -(1.f + e) * dvn * 0.5f;
            dvx += -J * nx;
            dvy += -J * ny;
            dvz += -J * nz;
        }
    }
};

// Pairs of the target at sorted position pos within `radius`: the search
// below the lowest cell that holds them all.
template <int CAP, class Pair>
__device__ inline void lane_search(const float4 *__restrict__ rc, const int4 *__restrict__ rb,
                                   const SearchTree &S, int pos, float tx, float ty, float tz,
                                   float radius, float guard, int2 *st, int *err, Pair &pair)
{
    const int j = search_start(S.opar, S.opar_leaf, S.bin_cell, pos, tx, ty, tz, radius, guard);
    if (j < 0) return;  // a disk of one body has no pairs
    lane_walk_block<CAP>(rc, rb, S, S.first[j], S.cnt[j], tx, ty, tz, radius, guard, st, err, pair);
}

// Collision deltas of wave w (64 consecutive sorted bodies); st: the lane's
// stack of CAP entries (stride kWave).
struct CollideArgs {
    DiskTable D;
    const float4 *rc;
    const float *vx, *vy, *vz;
    const Root *roots;
    float restitution;
    float *dv;
};

template <int CAP>
__device__ inline void collide_wave(const SegTable &T, const CollideArgs &C,
                                    const int4 *__restrict__ rb, const float4 *__restrict__ bodies,
                                    const int *__restrict__ perm, int *err, const SearchTree &S,
                                    int w, int2 *st)
{
    const int s = seg_of_wave(T, w);
    const int pos = T.off[s] + (w - T.wave[s]) * kWave + __lane_id();
    if (pos >= T.off[s + 1]) return;
    const float4 b = bodies[pos];
    const int i = perm[pos];
    const float r = C.D.d[s].r_coll;
    CollidePair v{b.x, b.y, b.z, r * r, C.restitution, C.vx[i], C.vy[i], C.vz[i], pos, perm,
                  C.vx, C.vy, C.vz, 0.f, 0.f, 0.f, 0};
    if (r > 0.f)
        lane_search<CAP>(C.rc, rb, S, pos, b.x, b.y, b.z, r, root_guard(C.roots[s]), st, err, v);
    const int n = T.n_total;
    C.dv[i] = v.dvx;
    C.dv[n + i] = v.dvy;
    C.dv[2 * n + i] = v.dvz;
    if (v.pairs) atomicAdd(err + 1, v.pairs);
}

__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_collide(SegTable T, CollideArgs C,
                                                        const int4 *__restrict__ rb,
                                                        const float4 *__restrict__ bodies,
                                                        const int *__restrict__ perm, int *err,
                                                        SearchTree S)
{
    __shared__ int2 st[kLStackC * kWalkBlock];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= T.wave[T.nseg]) return;
    collide_wave<kLStackC>(T, C, rb, bodies, perm, err, S, w,
                           st + wib * kLStackC * kWave + __lane_id());
}

// The step's walk and collision pass as one launch: blocks [0, ncb) compute
// the collision deltas (four waves of 64 bodies each, stacks of kLStackC
// entries in the walk's LDS), the rest are walk blocks.  The collision blocks
// come first, so they start first; most of their waves end within
// microseconds and the walk fills the compute units; no second stream and
// no cross-stream hand-off before or after the walk.
// Five waves per SIMD, as the walk alone gets from its LDS: the collision
// path is held to the register budget that allows it.
template <int TG, bool QUAD, int U, bool SELF>
__global__ __launch_bounds__(kGravBlock) __attribute__((amdgpu_waves_per_eu(5))) void
k_walk_collide(SegTable T, WaveMap M, GravArgs A, int ncb, CollideArgs C, SearchTree S)
{
    __shared__ GravLds<TG, QUAD> lds[kGravWaves];
    static_assert(sizeof(lds) >= sizeof(int2) * kLStackC * kGravBlock,
                  "the collision stacks fit in the walk's LDS");
    if ((int)blockIdx.x < ncb) {  // each collision wave takes every (ncb * 4)-th body wave
        const int wib = threadIdx.x >> 6;
        int2 *st = reinterpret_cast<int2 *>(lds) + wib * kLStackC * kWave + __lane_id();
        for (int w = blockIdx.x * kGravWaves + wib; w < T.wave[T.nseg]; w += ncb * kGravWaves)
            collide_wave<kLStackC>(T, C, A.rb, A.bodies, A.perm, A.err, S, w, st);
        return;
    }
    walk_block<TG, QUAD, U, SELF>(T, M, A, blockIdx.x - ncb, lds);
}

// ---------------------------------------------------------------------------
// FoF: the predicate of Aggregate.cpp:38-40, (dx^2 + dy^2) + dz^2 < ll^2 in
// float without contraction; pairs are merged by linking the larger root
// under the smaller one, so each group's root is its smallest index (the
// canonical label).
// ---------------------------------------------------------------------------
__device__ inline int uf_find(int *parent, int x)
{
    int p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        const int g = __hip_atomic_load(&parent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g != p)  // path halving; only ever moves a pointer toward its root
            __hip_atomic_store(&parent[x], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = g;
    }
    return x;
}

__device__ inline void uf_union(int *parent, int a, int b)
{
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        // a > b: hook root a under b if a is still a root
        int expected = a;
        if (__hip_atomic_compare_exchange_strong(&parent[a], &expected, b, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

struct FofPair {
    float bx, by, bz, ll2;
    int self_i;
    const int *perm;
    int *parent;
    __device__ inline void operator()(const float4 &c, int p)
    {
#pragma clang fp contract(off)
        const int j = perm[p];
        if (j <= self_i) return;
        // Aggregate.cpp:38 computes p[i]-p[j] with i < j
        const float dx = bx - c.x, dy = by - c.y, dz = bz - c.z;
        if (dx * dx + dy * dy + dz * dz < ll2) uf_union(parent, self_i, j);
    }
};

__global__ __launch_bounds__(kWalkBlock) void k_fof_link(SegTable T, const float4 *__restrict__ rc,
                                                     const int4 *__restrict__ rb,
                                                     const float4 *__restrict__ bodies,
                                                     const int *__restrict__ perm,
                                                     const Root *__restrict__ roots, LinkTable LT,
                                                     int *parent, int *err, SearchTree S)
{
    __shared__ int2 st[kLStack * kWalkBlock];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= T.wave[T.nseg]) return;
    const int s = seg_of_wave(T, w);
    const int pos = T.off[s] + (w - T.wave[s]) * kWave + __lane_id();
    if (pos >= T.off[s + 1]) return;
    const float4 b = bodies[pos];
    const float ll = LT.ll[s];  // the disk's own link length
    FofPair v{b.x, b.y, b.z, ll * ll, perm[pos], perm, parent};
    if (ll > 0.f)
        lane_search<kLStack>(rc, rb, S, pos, b.x, b.y, b.z, ll, root_guard(roots[s]),
                    st + wib * kLStack * kWave + __lane_id(), err, v);
}

__global__ void k_iota(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Identity order in both parity halves of the wave-order buffer.
__global__ void k_iota2(int *p, int stride)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * stride) p[i] = i < stride ? i : i - stride;
}

__global__ void k_fof_final(int *parent, int *labels, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    labels[i] = uf_find(parent, i);
}

// ---------------------------------------------------------------------------
// FP32 probes: the achievable rates the walk is compared with.
// ---------------------------------------------------------------------------
constexpr int kProbeBlock = 256;

__global__ __launch_bounds__(kProbeBlock) void k_probe_fma(float *out, int iters, float a, float b)
{
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (float)(threadIdx.x + k) * 1e-3f;
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = fmaf(v[k], a, b);
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += v[k];
    if (sum == 1234.5f) out[blockIdx.x] = sum;  // keeps the chains live
}

// The walk's record evaluation (GravityVisitor::visit_masked) for one target
// per lane over a table of 64 records read with LDS broadcasts.
__global__ __launch_bounds__(kProbeBlock) void k_probe_mix(float *out, int iters, float theta2,
                                                           float eps2)
{
    __shared__ float4 ta[64];
    __shared__ float th2[64];
    const int t = threadIdx.x;
    if (t < 64) {
        const float u = (float)t;
        ta[t] = make_float4(0.7f * u - 20.f, 0.05f * u, 0.3f * u - 9.f, 1e-6f * (u + 1.f));
        th2[t] = (t & 3) == 0 ? -1.f : 0.01f * u / theta2;  // staged h^2 theta^-2
    }
    __syncthreads();
    const float bx = 0.01f * t, by = 0.001f * t, bz = -0.01f * t;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int i = 0; i < iters; ++i) {
        const int k = i & 63;
        const float4 a = ta[k];
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float r2 = dist2(dx, dy, dz);
        const bool open = !(th2[k] < r2);
        const bool take = !open && k != (t & 63);
        const float rinv = __builtin_amdgcn_rsqf(r2 + eps2);
        const float mw = take ? a.w : 0.f;
        const float f = mw * (rinv * rinv * rinv);
        fx += f * dx;
        fy += f * dy;
        fz += f * dz;
    }
    if (fx + fy + fz == 1234.5f) out[blockIdx.x] = fx;
}

double probe_fp32(int kind, float *scratch, hipStream_t st)
{
    const int blocks = 256 * 4 * 8;  // 8 waves per SIMD across the chip
    const int iters = kind == 0 ? 1 << 14 : 1 << 13;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0;
    auto launch = [&]() {
        if (kind == 0)
            hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(kProbeBlock), 0, st, scratch,
                               iters, 0.999f, 1e-3f);
        else
            hipLaunchKernelGGL(k_probe_mix, dim3(blocks), dim3(kProbeBlock), 0, st, scratch,
                               iters, 0.25f, 1e-6f);
    };
    launch();  // warm-up
    (void)hipEventRecord(e0, st);
    const int reps = 3;
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double per_iter = kind == 0 ? 16.0 : 20.0;  // flop per thread and iteration
    const double flop = per_iter * iters * (double)blocks * kProbeBlock * reps;
    return ms > 0.f ? flop / (ms * 1e-3) / 1e12 : -1.0;
}

// ===========================================================================
// Launchers
// ===========================================================================

constexpr int kUnroll64 = 2;

static CollideArgs collide_args(const DiskTable &D, const float *vx, const float *vy,
                               const float *vz, float restitution, Bufs &B)
{
    return CollideArgs{D, B.rec_c, vx, vy, vz, B.roots, restitution, B.dv};
}

template <int TG>
static void launch_walk_tg(const SegTable &T, const DiskDev *dtab, const int *step_ctr,
                           const WalkArgs &W, Bufs &B, hipStream_t st, const CollideArgs *C)
{
    WaveMap M;
    M.wave[0] = 0;
    for (int s = 0; s < kMaxSeg; ++s)
        M.wave[s + 1] = M.wave[s] + (s < T.nseg ? cdiv(T.off[s + 1] - T.off[s], TG) : 0);
    const int waves = M.wave[T.nseg];
    // with one target per lane a lane evaluates kUnroll64 records per step
    constexpr int U = TG >= 16 ? kUnroll64 : 1;
    constexpr int UQ = 1;  // quadrupole records: one per lane and step
    const bool lpt = W.lpt && step_ctr && !W.xcd;
    const int chunk = W.xcd ? cdiv(waves, kXcds) : 0;
    const int blocks = W.xcd ? kXcds * cdiv(chunk, kGravWaves) : cdiv(waves, kGravWaves);
    GravArgs A{dtab, step_ctr, B.rec_a, B.rec_b, W.quad ? B.rec_q : nullptr, B.bodies,
               B.vals[B.cur], W.theta2, W.eps2, W.softening, W.star_mass, W.ax, W.ay, W.az,
               W.sax, W.say, W.saz, B.err, lpt ? B.lpt_order : nullptr,
               lpt ? B.lpt_cost : nullptr, B.lpt_stride, chunk};
    // the own-leaf test is needed only without softening
    const bool self = W.eps2 < kNoSelfTestEps2;
    if (!C) {
        auto kern = !self ? (W.quad ? k_walk<TG, true, UQ, false> : k_walk<TG, false, U, false>)
                          : (W.quad ? k_walk<TG, true, UQ, true> : k_walk<TG, false, U, true>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kGravBlock), 0, st, T, M, A);
        return;
    }
    // collision blocks: a multiple of the XCD count, so the walk blocks keep
    // their die assignment
    const int ncb = kXcds * cdiv(cdiv(T.wave[T.nseg], kGravWaves * kCollideRounds), kXcds);
    auto kern = !self ? (W.quad ? k_walk_collide<TG, true, UQ, false>
                                : k_walk_collide<TG, false, U, false>)
                      : (W.quad ? k_walk_collide<TG, true, UQ, true>
                                : k_walk_collide<TG, false, U, true>);
    hipLaunchKernelGGL(kern, dim3(ncb + blocks), dim3(kGravBlock), 0, st, T, M, A, ncb, *C,
                       search_bufs(B));
}

// ---------------------------------------------------------------------------
// Wave order of the next walk: a counting sort of the wave costs of the last
// walk, longest first.  Step parity p: the walk of a step reads order[p] and
// writes cost[p]; this pass, run beside the tree build of the same step,
// reads cost[p ^ 1] (the previous walk) and writes order[p ^ 1] (the next
// walk).  Any permutation is correct; the order only shortens the tail of
// the launch.
// ---------------------------------------------------------------------------
void reset_lpt(Bufs &B, hipStream_t st)
{
    const int n = 2 * B.lpt_stride;
    hipLaunchKernelGGL(k_iota2, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, B.lpt_order,
                       B.lpt_stride);
    launch_fill_int(B.lpt_cost, 0, n, st);
}

LptDev lpt_dev(Bufs &B, int nw, const int *step_ctr)
{
    LptDev L;
    L.step_ctr = step_ctr;
    L.cost = B.lpt_cost;
    L.order = B.lpt_order;
    L.hist = B.lpt_hist;
    L.stride = B.lpt_stride;
    L.nw = nw;
    return L;
}

static void launch_walk_any(const SegTable &T, const DiskDev *dtab, const int *step_ctr,
                            const WalkArgs &W, Bufs &B, hipStream_t st, const CollideArgs *C)
{
    switch (W.targets) {
    case 8: launch_walk_tg<8>(T, dtab, step_ctr, W, B, st, C); break;
    case 16: launch_walk_tg<16>(T, dtab, step_ctr, W, B, st, C); break;
    case 32: launch_walk_tg<32>(T, dtab, step_ctr, W, B, st, C); break;
    default: launch_walk_tg<64>(T, dtab, step_ctr, W, B, st, C); break;
    }
}

void launch_walk(const SegTable &T, const DiskDev *dtab, const int *step_ctr, const WalkArgs &W,
                 Bufs &B, hipStream_t st)
{
    launch_walk_any(T, dtab, step_ctr, W, B, st, nullptr);
}

void launch_walk_collide(const SegTable &T, const DiskDev *dtab, const int *step_ctr,
                         const WalkArgs &W, const DiskTable &D, const float *vx, const float *vy,
                         const float *vz, float restitution, Bufs &B, hipStream_t st)
{
    const CollideArgs C = collide_args(D, vx, vy, vz, restitution, B);
    launch_walk_any(T, dtab, step_ctr, W, B, st, &C);
}

template <int TG>
static void launch_counts_tg(const SegTable &T, float theta2, int *n_acc, int *n_open,
                             int *wstat, Bufs &B, hipStream_t st)
{
    WaveMap M;
    M.wave[0] = 0;
    for (int s = 0; s < kMaxSeg; ++s)
        M.wave[s + 1] = M.wave[s] + (s < T.nseg ? cdiv(T.off[s + 1] - T.off[s], TG) : 0);
    hipLaunchKernelGGL(k_walk_counts<TG>, dim3(cdiv(M.wave[T.nseg], kGravWaves)),
                       dim3(kGravBlock), 0, st, T, M, B.rec_a, B.rec_b, B.bodies, B.vals[B.cur],
                       theta2, n_acc, n_open, B.err, wstat);
}

int walk_waves(const SegTable &T, int targets)
{
    int w = 0;
    for (int s = 0; s < T.nseg; ++s) w += cdiv(T.off[s + 1] - T.off[s], targets);
    return w;
}

void launch_walk_counts(const SegTable &T, float theta2, int targets, int *n_acc, int *n_open,
                        int *wstat, Bufs &B, hipStream_t st)
{
    switch (targets) {
    case 8: launch_counts_tg<8>(T, theta2, n_acc, n_open, wstat, B, st); break;
    case 16: launch_counts_tg<16>(T, theta2, n_acc, n_open, wstat, B, st); break;
    case 32: launch_counts_tg<32>(T, theta2, n_acc, n_open, wstat, B, st); break;
    default: launch_counts_tg<64>(T, theta2, n_acc, n_open, wstat, B, st); break;
    }
}

void launch_collide(const SegTable &T, const DiskTable &D, const float *vx, const float *vy,
                    const float *vz, float restitution, Bufs &B, hipStream_t st)
{
    const int waves = T.wave[T.nseg];
    hipLaunchKernelGGL(k_collide, dim3(cdiv(waves, kWalkWaves)), dim3(kWalkBlock), 0, st, T,
                       collide_args(D, vx, vy, vz, restitution, B), B.rec_b, B.bodies,
                       B.vals[B.cur], B.err, search_bufs(B));
}

void launch_fof(const SegTable &T, const LinkTable &LT, int *labels, Bufs &B, hipStream_t st)
{
    const int n = T.n_total, waves = T.wave[T.nseg];
    hipLaunchKernelGGL(k_iota, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, B.labels, n);
    hipLaunchKernelGGL(k_fof_link, dim3(cdiv(waves, kWalkWaves)), dim3(kWalkBlock), 0, st, T,
                       B.rec_c, B.rec_b, B.bodies, B.vals[B.cur], B.roots, LT, B.labels, B.err,
                       search_bufs(B));
    hipLaunchKernelGGL(k_fof_final, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, B.labels,
                       labels, n);
}

}  // namespace nbx
