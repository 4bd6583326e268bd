// nbx_walk.hip -- tree walks on gfx950: gravity, collision impulse, FoF.
//
// All three walks share one traversal: a wavefront owns 64 consecutive
// targets in Morton order (one per lane) and walks the compressed octree
// with a wave-uniform stack kept in LDS.  A stack entry is a child block
// {first record, count} plus the 64-bit mask of the lanes whose own walk
// opened the parent.  Records are read with wave-uniform addresses (scalar
// loads through the constant cache); each lane evaluates a record only if
// its bit is set, so every target visits exactly the nodes its own
// per-body walk (BHTree.cu:555-593) visits, and the wave shares the loads.
#include "nbx_device.hpp"
#include "nbx_launch.hpp"

namespace nbx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

constexpr int kWalkWaves = kBlock / kWave;  // waves per block

struct StackEntry {
    int first, count;
    unsigned mlo, mhi;
};

// Visitor interface: visit(a, b, leaf, mine) for one record; returns true
// when the lane wants the record's children (internal records only).
template <class Visitor>
__device__ inline void wave_walk(const float4 *__restrict__ ra, const int4 *__restrict__ rb,
                                 int root, uint64_t active, StackEntry *st, int *err,
                                 Visitor &vis)
{
    const int lane = __lane_id();
    int sp = 0;
    // every lane writes the same entry: per-lane program order covers the
    // later reads, no cross-lane ordering is needed
    st[0] = StackEntry{root, 1, (unsigned)active, (unsigned)(active >> 32)};
    sp = 1;
    while (sp > 0) {
        --sp;
        const StackEntry e = st[sp];
        const int first = rfl(e.first);
        const int count = rfl(e.count);
        const uint64_t mask = ((uint64_t)(unsigned)rfl((int)e.mhi) << 32) |
                              (uint64_t)(unsigned)rfl((int)e.mlo);
        const bool mine = (mask >> lane) & 1ull;
        if (count <= 8) {
            float4 A[8];
            int4 Bv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                A[k] = ra[first + k];
                Bv[k] = rb[first + k];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k < count) {
                    const bool leaf = Bv[k].w == 0;
                    const bool open = mine && vis.visit(A[k], Bv[k], leaf);
                    const uint64_t om = ballot64(open);
                    if (om) {
                        if (sp < kStackCap) {
                            st[sp] = StackEntry{Bv[k].z, Bv[k].w, (unsigned)om,
                                                (unsigned)(om >> 32)};
                            ++sp;
                        } else if (lane == 0) {
                            atomicOr(err, kErrStack);
                        }
                    }
                }
            }
        } else {
            for (int k = 0; k < count; ++k) {
                const float4 a = ra[first + k];
                const int4 b = rb[first + k];
                const bool leaf = b.w == 0;
                const bool open = mine && vis.visit(a, b, leaf);
                const uint64_t om = ballot64(open);
                if (om) {
                    if (sp < kStackCap) {
                        st[sp] = StackEntry{b.z, b.w, (unsigned)om, (unsigned)(om >> 32)};
                        ++sp;
                    } else if (lane == 0) {
                        atomicOr(err, kErrStack);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Gravity: accept a record when it is a leaf or h^2 < theta^2 r^2 (the
// half-width / COM-distance test of BHTree.cu:576-577); skip the target's
// own leaf (BHTree.cu:580); monopole pull m d / (r^2+eps^2)^{3/2}.
// ---------------------------------------------------------------------------
struct GravityVisitor {
    float bx, by, bz, theta2, eps2;
    int self;
    float fx, fy, fz;
    __device__ inline bool visit(const float4 &a, const int4 &b, bool leaf)
    {
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool open = !leaf && !(__int_as_float(b.y) < theta2 * r2);
        if (!open && !(leaf && b.z == self)) {
            const float rinv = __builtin_amdgcn_rsqf(r2 + eps2);
            const float f = a.w * rinv * rinv * rinv;
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
        }
        return open;
    }
};

__global__ __launch_bounds__(kBlock) void k_walk(
    SegTable T, DiskTable D, const float4 *__restrict__ ra, const int4 *__restrict__ rb,
    const float4 *__restrict__ bodies, const int *__restrict__ perm, float theta2, float eps2,
    float softening, float star_mass, float *ax, float *ay, float *az, float *sax, float *say,
    float *saz, int *err)
{
    __shared__ StackEntry stk[kWalkWaves][kStackCap];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= T.wave[T.nseg]) return;  // uniform per wave
    const int s = seg_of_wave(T, w);
    const int pos = T.off[s] + (w - T.wave[s]) * kWave + __lane_id();
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    GravityVisitor v{b.x, b.y, b.z, theta2, eps2, pos, 0.f, 0.f, 0.f};
    wave_walk(ra, rb, T.noff[s], ballot64(valid), stk[wib], err, v);
    if (!valid) return;
    const int i = perm[pos];
    float ox, oy, oz;
    {
#pragma clang fp contract(off)
        // star at the origin (BHTree.cu:547-553)
        float dx = -b.x, dy = -b.y, dz = -b.z;
        float r2 = dx * dx + dy * dy + dz * dz + softening * softening;
        float r = sqrtf(r2);
        float f = star_mass / (r2 * r);
        ox = f * dx + v.fx;
        oy = f * dy + v.fy;
        oz = f * dz + v.fz;
        const DiskDev &d = D.d[s];
        if (d.pmass > 0.f)
            perturber_add_dev(b.x, b.y, b.z, d.px, d.py, d.pz, d.pmass, softening, ox, oy, oz);
    }
    ax[i] = ox;
    ay[i] = oy;
    az[i] = oz;
    if (sax) {
        sax[i] = v.fx;
        say[i] = v.fy;
        saz[i] = v.fz;
    }
}

// Interaction counts of the same walk (accepted records excluding the own
// leaf, opened internal records): the algorithmic work of the force pass.
struct CountVisitor {
    float bx, by, bz, theta2;
    int self;
    int n_acc, n_open;
    __device__ inline bool visit(const float4 &a, const int4 &b, bool leaf)
    {
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool open = !leaf && !(__int_as_float(b.y) < theta2 * r2);
        if (open) ++n_open;
        else if (!(leaf && b.z == self)) ++n_acc;
        return open;
    }
};

__global__ __launch_bounds__(kBlock) void k_walk_counts(SegTable T, const float4 *__restrict__ ra,
                                                        const int4 *__restrict__ rb,
                                                        const float4 *__restrict__ bodies,
                                                        const int *__restrict__ perm,
                                                        float theta2, int *n_acc, int *n_open,
                                                        int *err)
{
    __shared__ StackEntry stk[kWalkWaves][kStackCap];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= T.wave[T.nseg]) return;
    const int s = seg_of_wave(T, w);
    const int pos = T.off[s] + (w - T.wave[s]) * kWave + __lane_id();
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    CountVisitor v{b.x, b.y, b.z, theta2, pos, 0, 0};
    wave_walk(ra, rb, T.noff[s], ballot64(valid), stk[wib], err, v);
    if (!valid) return;
    const int i = perm[pos];
    n_acc[i] = v.n_acc;
    n_open[i] = v.n_open;
}

// ---------------------------------------------------------------------------
// Collision (BHTree.cu:620-673) with an exact pair set: a cell is opened
// when it may hold a body within r_coll (every body of a cell of half-size
// h lies within 2 sqrt(3) h of its COM); leaves test dist^2 < r_coll^2.
// ---------------------------------------------------------------------------
struct CollideVisitor {
    float bx, by, bz, rc, rc2, e;
    float vix, viy, viz;
    int self;
    const int *perm;
    const float *vx, *vy, *vz;
    float dvx, dvy, dvz;
    int pairs;
    __device__ inline bool visit(const float4 &a, const int4 &b, bool leaf)
    {
#pragma clang fp contract(off)
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float dist2 = dx * dx + dy * dy + dz * dz;
        if (!leaf) return sqrtf(dist2) - 3.5f * __int_as_float(b.x) < rc * 1.01f;
        if (b.z == self || !(dist2 < rc2)) return false;
        ++pairs;
        const int j = perm[b.z];
        const float dist = sqrtf(dist2) + 1e-10f;
        const float nx = dx / dist, ny = dy / dist, nz = dz / dist;
        const float dvn = (vx[j] - vix) * nx + (vy[j] - viy) * ny + (vz[j] - viz) * nz;
        if (dvn < 0.f) {
            const float J = -(1.f + e) * dvn * 0.5f;
            dvx += -J * nx;
            dvy += -J * ny;
            dvz += -J * nz;
        }
        return false;
    }
};

__global__ __launch_bounds__(kBlock) void k_collide(
    SegTable T, DiskTable D, const float4 *__restrict__ ra, const int4 *__restrict__ rb,
    const float4 *__restrict__ bodies, const int *__restrict__ perm,
    const float *__restrict__ vx, const float *__restrict__ vy, const float *__restrict__ vz,
    float restitution, float *dv, int *err)
{
    __shared__ StackEntry stk[kWalkWaves][kStackCap];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= T.wave[T.nseg]) return;
    const int s = seg_of_wave(T, w);
    const int pos = T.off[s] + (w - T.wave[s]) * kWave + __lane_id();
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    const int i = perm[valid ? pos : T.off[s]];
    const float rc = D.d[s].r_coll;
    CollideVisitor v{b.x, b.y, b.z, rc, rc * rc, restitution, vx[i], vy[i], vz[i], pos, perm,
                     vx, vy, vz, 0.f, 0.f, 0.f, 0};
    wave_walk(ra, rb, T.noff[s], ballot64(valid), stk[wib], err, v);
    if (!valid) return;
    const int n = T.n_total;
    dv[i] = v.dvx;
    dv[n + i] = v.dvy;
    dv[2 * n + i] = v.dvz;
    if (v.pairs) atomicAdd(err + 1, v.pairs);
}

// ---------------------------------------------------------------------------
// FoF: same neighbour walk with the predicate of Aggregate.cpp:38-40,
// (dx^2 + dy^2) + dz^2 < ll^2 in float without contraction; pairs are
// merged by linking the larger root under the smaller one, so each group's
// root is its smallest index (the canonical label).
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

struct FofVisitor {
    float bx, by, bz, ll, ll2;
    int self_i;
    const int *perm;
    int *parent;
    __device__ inline bool visit(const float4 &a, const int4 &b, bool leaf)
    {
#pragma clang fp contract(off)
        if (!leaf) {
            const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
            const float d2 = dx * dx + dy * dy + dz * dz;
            return sqrtf(d2) - 3.5f * __int_as_float(b.x) < ll * 1.01f;
        }
        const int j = perm[b.z];
        if (j <= self_i) return false;
        // Aggregate.cpp:38 computes p[i]-p[j] with i < j
        const float dx = bx - a.x, dy = by - a.y, dz = bz - a.z;
        if (dx * dx + dy * dy + dz * dz < ll2) uf_union(parent, self_i, j);
        return false;
    }
};

__global__ __launch_bounds__(kBlock) void k_fof_link(SegTable T, const float4 *__restrict__ ra,
                                                     const int4 *__restrict__ rb,
                                                     const float4 *__restrict__ bodies,
                                                     const int *__restrict__ perm, float ll,
                                                     int *parent, int *err)
{
    __shared__ StackEntry stk[kWalkWaves][kStackCap];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= T.wave[T.nseg]) return;
    const int s = seg_of_wave(T, w);
    const int pos = T.off[s] + (w - T.wave[s]) * kWave + __lane_id();
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    FofVisitor v{b.x, b.y, b.z, ll, ll * ll, perm[valid ? pos : T.off[s]], perm, parent};
    wave_walk(ra, rb, T.noff[s], ballot64(valid), stk[wib], err, v);
}

__global__ void k_iota(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_fof_final(int *parent, int *labels, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    labels[i] = uf_find(parent, i);
}

// ===========================================================================
// Launchers
// ===========================================================================

void launch_walk(const SegTable &T, const DiskTable &D, const WalkArgs &W, Bufs &B,
                 hipStream_t st)
{
    const int waves = T.wave[T.nseg];
    hipLaunchKernelGGL(k_walk, dim3(cdiv(waves, kWalkWaves)), dim3(kBlock), 0, st, T, D,
                       B.rec_a, B.rec_b, B.bodies, B.vals[0], W.theta2, W.eps2, W.softening,
                       W.star_mass, W.ax, W.ay, W.az, W.sax, W.say, W.saz, B.err);
}

void launch_walk_counts(const SegTable &T, float theta2, int *n_acc, int *n_open, Bufs &B,
                        hipStream_t st)
{
    const int waves = T.wave[T.nseg];
    hipLaunchKernelGGL(k_walk_counts, dim3(cdiv(waves, kWalkWaves)), dim3(kBlock), 0, st, T,
                       B.rec_a, B.rec_b, B.bodies, B.vals[0], theta2, n_acc, n_open, B.err);
}

void launch_collide(const SegTable &T, const DiskTable &D, const float *vx, const float *vy,
                    const float *vz, float restitution, Bufs &B, hipStream_t st)
{
    const int waves = T.wave[T.nseg];
    hipLaunchKernelGGL(k_collide, dim3(cdiv(waves, kWalkWaves)), dim3(kBlock), 0, st, T, D,
                       B.rec_a, B.rec_b, B.bodies, B.vals[0], vx, vy, vz, restitution, B.dv,
                       B.err);
}

void launch_fof(const SegTable &T, float ll, int *labels, Bufs &B, hipStream_t st)
{
    const int n = T.n_total, waves = T.wave[T.nseg];
    hipLaunchKernelGGL(k_iota, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, B.labels, n);
    hipLaunchKernelGGL(k_fof_link, dim3(cdiv(waves, kWalkWaves)), dim3(kBlock), 0, st, T,
                       B.rec_a, B.rec_b, B.bodies, B.vals[0], ll, B.labels, B.err);
    hipLaunchKernelGGL(k_fof_final, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, B.labels,
                       labels, n);
}

}  // namespace nbx
