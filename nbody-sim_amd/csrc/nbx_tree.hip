// nbx_tree.hip -- octree construction and integrator kernels for gfx950.
//
// Build pipeline (one launch per phase for all disks of a batch):
//   bbox partials -> root cube -> 63-bit keys -> LSD radix sort (8 x 8-bit)
//   -> gather sorted bodies -> binary radix tree (one thread per internal
//   node, longest-common-prefix search) -> COM bottom-up (second arrival
//   climbs) -> compressed octree records (children contiguous).
// Every phase is HBM-streaming integer/float work; nothing here is a GEMM.
#include <float.h>

#include "nbx_device.hpp"
#include "nbx_launch.hpp"

namespace nbx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ===========================================================================
// Integrator kernels.  Float expressions are written exactly as in
// Integrator.cu / Perturber.cu and compiled without contraction so the
// results round identically to the reference's arithmetic.
// ===========================================================================

__global__ void k_half_kick(float *vx, float *vy, float *vz, const float *__restrict__ ax,
                            const float *__restrict__ ay, const float *__restrict__ az,
                            float dt_half, int n)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] += ax[i] * dt_half;  // Integrator.cu:12-14
    vy[i] += ay[i] * dt_half;
    vz[i] += az[i] * dt_half;
}

__global__ void k_drift(float *x, float *y, float *z, const float *__restrict__ vx,
                        const float *__restrict__ vy, const float *__restrict__ vz, float dt,
                        int n)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += vx[i] * dt;  // Integrator.cu:22-24
    y[i] += vy[i] * dt;
    z[i] += vz[i] * dt;
}

__global__ void k_reset_acc(float *ax, float *ay, float *az, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ax[i] = 0.f;
    ay[i] = 0.f;
    az[i] = 0.f;
}

__global__ void k_gas_drag(const float *__restrict__ x, const float *__restrict__ z, float *vx,
                           float *vy, float *vz, float decay, float star_mass, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = vx[i], b = vy[i], c = vz[i];
    gas_drag_one(x[i], z[i], a, b, c, decay, star_mass);
    vx[i] = a;
    vy[i] = b;
    vz[i] = c;
}

__global__ void k_perturber(float *ax, float *ay, float *az, const float *__restrict__ x,
                            const float *__restrict__ y, const float *__restrict__ z, float px,
                            float py, float pz, float pmass, float soft, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = ax[i], b = ay[i], c = az[i];
    perturber_add_dev(x[i], y[i], z[i], px, py, pz, pmass, soft, a, b, c);
    ax[i] = a;
    ay[i] = b;
    az[i] = c;
}

// Fused first half-kick + drift (main.cu:104-107) with the bbox partial of
// the drifted positions (BHTree.cu:94-133) in the same pass over HBM.
__global__ __launch_bounds__(kBlock) void k_kick_drift_bbox(
    SegTable T, float *x, float *y, float *z, float *vx, float *vy, float *vz,
    const float *__restrict__ ax, const float *__restrict__ ay, const float *__restrict__ az,
    float dt_half, float dt, float *partial)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    const int s = seg_of_block(T, b);
    const int i = T.off[s] + (b - T.blk[s]) * kBlock + threadIdx.x;
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
    float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    if (i < T.off[s + 1]) {
        float vxi = vx[i] + ax[i] * dt_half;
        float vyi = vy[i] + ay[i] * dt_half;
        float vzi = vz[i] + az[i] * dt_half;
        float xi = x[i] + vxi * dt;
        float yi = y[i] + vyi * dt;
        float zi = z[i] + vzi * dt;
        vx[i] = vxi; vy[i] = vyi; vz[i] = vzi;
        x[i] = xi; y[i] = yi; z[i] = zi;
        lx = hx = xi;
        ly = hy = yi;
        lz = hz = zi;
    }
    block_minmax6(lx, ly, lz, hx, hy, hz, partial + 6 * b);
}

// Collision delta, second half-kick and gas drag (main.cu:131-139 after the
// impulse pass): v += dv; v += a*dt_half; drag.
__global__ void k_kick_drag(float *vx, float *vy, float *vz, const float *__restrict__ ax,
                            const float *__restrict__ ay, const float *__restrict__ az,
                            const float *__restrict__ dv, const float *__restrict__ x,
                            const float *__restrict__ z, float dt_half, int has_drag,
                            float decay, float star_mass, int n)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = vx[i], b = vy[i], c = vz[i];
    if (dv) {
        a += dv[i];
        b += dv[n + i];
        c += dv[2 * n + i];
    }
    if (dt_half != 0.f) {
        a += ax[i] * dt_half;
        b += ay[i] * dt_half;
        c += az[i] * dt_half;
    }
    if (has_drag) gas_drag_one(x[i], z[i], a, b, c, decay, star_mass);
    vx[i] = a;
    vy[i] = b;
    vz[i] = c;
}

__global__ void k_fill_int(int *p, int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_float(float *p, float v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ===========================================================================
// Bounding box and root cube
// ===========================================================================

__global__ __launch_bounds__(kBlock) void k_bbox(SegTable T, const float *__restrict__ x,
                                                 const float *__restrict__ y,
                                                 const float *__restrict__ z, float *partial)
{
    const int b = blockIdx.x;
    const int s = seg_of_block(T, b);
    const int i = T.off[s] + (b - T.blk[s]) * kBlock + threadIdx.x;
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
    float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    if (i < T.off[s + 1]) {
        lx = hx = x[i];
        ly = hy = y[i];
        lz = hz = z[i];
    }
    block_minmax6(lx, ly, lz, hx, hy, hz, partial + 6 * b);
}

// One block per disk: reduce the partials, pad by 1.001 about the midpoint
// (BHTree.cu:440-446) and form the root cube (BHTree.cu:187-191).
__global__ __launch_bounds__(kBlock) void k_root(SegTable T, const float *__restrict__ partial,
                                                 Root *roots)
{
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
    float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    for (int b = T.blk[s] + threadIdx.x; b < T.blk[s + 1]; b += kBlock) {
        const float *p = partial + 6 * b;
        lx = fminf(lx, p[0]); ly = fminf(ly, p[1]); lz = fminf(lz, p[2]);
        hx = fmaxf(hx, p[3]); hy = fmaxf(hy, p[4]); hz = fmaxf(hz, p[5]);
    }
    __shared__ float r[6];
    block_minmax6(lx, ly, lz, hx, hy, hz, r);
    __syncthreads();
    if (threadIdx.x == 0) {
        float bmin[3], bmax[3];
        for (int k = 0; k < 3; ++k) {
            float lo = r[k], hi = r[3 + k];
            float mid = (lo + hi) * 0.5f;
            float half = (hi - lo) * 0.5f * 1.001f;
            bmin[k] = mid - half;
            bmax[k] = mid + half;
        }
        Root R;
        R.c[0] = (bmin[0] + bmax[0]) * 0.5f;
        R.c[1] = (bmin[1] + bmax[1]) * 0.5f;
        R.c[2] = (bmin[2] + bmax[2]) * 0.5f;
        R.c[3] = fmaxf(fmaxf(bmax[0] - bmin[0], bmax[1] - bmin[1]), bmax[2] - bmin[2]) * 0.5f *
                 1.001f;
        roots[s] = R;
    }
}

// ===========================================================================
// Keys: the octant digits of BHTree.cu:164-170 along the float descent of
// BHTree.cu:253-256, 21 levels, first level most significant.
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_keys(SegTable T, const float *__restrict__ x,
                                                 const float *__restrict__ y,
                                                 const float *__restrict__ z,
                                                 const Root *__restrict__ roots,
                                                 uint64_t *keys, int *vals)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    const int s = seg_of_block(T, b);
    const int i = T.off[s] + (b - T.blk[s]) * kBlock + threadIdx.x;
    if (i >= T.off[s + 1]) return;
    const Root R = roots[s];
    float cx = R.c[0], cy = R.c[1], cz = R.c[2], h = R.c[3];
    const float px = x[i], py = y[i], pz = z[i];
    uint64_t key = 0;
#pragma unroll
    for (int l = 0; l < kLevels; ++l) {
        const bool ox = px > cx, oy = py > cy, oz = pz > cz;
        key = (key << 3) | (uint64_t)((ox ? 1 : 0) | (oy ? 2 : 0) | (oz ? 4 : 0));
        h = h * 0.5f;
        cx = cx + (ox ? h : -h);
        cy = cy + (oy ? h : -h);
        cz = cz + (oz ? h : -h);
    }
    keys[i] = key;
    vals[i] = i;
}

// ===========================================================================
// LSD radix sort, 8-bit digits, stable.  Per pass: tile histograms
// (digit-major), one exclusive scan, stable scatter (wave match + LDS).
// mode 0: digit = (key >> shift) & 255;  mode 1: digit = disk of the body.
// ===========================================================================

__device__ inline int sort_digit(int mode, uint64_t key, int val, int shift, const SegTable &T)
{
    return mode == 0 ? (int)((key >> shift) & 255u) : seg_of_body(T, val);
}

__global__ __launch_bounds__(kBlock) void k_sort_hist(const uint64_t *__restrict__ keys,
                                                      const int *__restrict__ vals, int n,
                                                      int shift, int mode, SegTable T,
                                                      int n_tiles, int *hist)
{
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        int i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[sort_digit(mode, keys[i], vals[i], shift, T)], 1);
    }
    __syncthreads();
    hist[threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Single-block exclusive scan in place, chunks of 4 * 1024 with a carry.
__global__ __launch_bounds__(1024) void k_scan1(int *data, int n)
{
    __shared__ int wsum[16];
    __shared__ int carry_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 4096) {
        int v[4];
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int i = base + 4 * t + k;
            v[k] = i < n ? data[i] : 0;
            sum += v[k];
        }
        int inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (t < 16) {
            int x = wsum[t];
            int acc = x;
            for (int o = 1; o < 16; o <<= 1) {
                int y = __shfl_up(acc, o, 16);
                if (t >= o) acc += y;
            }
            wsum[t] = acc - x;  // exclusive prefix of the wave sums
        }
        __syncthreads();
        int run = carry_s + wsum[w] + inc - sum;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int i = base + 4 * t + k;
            if (i < n) data[i] = run;
            run += v[k];
        }
        __syncthreads();
        if (t == 1023) carry_s = run;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(
    const uint64_t *__restrict__ keys_in, const int *__restrict__ vals_in, uint64_t *keys_out,
    int *vals_out, int n, int shift, int mode, SegTable T, int n_tiles,
    const int *__restrict__ hist)
{
    __shared__ int s_off[256];
    __shared__ int s_wc[kBlock / kWave][256];
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    s_off[t] = hist[t * n_tiles + blockIdx.x];
    const int base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
#pragma unroll
        for (int k = 0; k < kBlock / kWave; ++k) s_wc[k][t] = 0;
        __syncthreads();
        const int i = base + r * kBlock + t;
        const bool valid = i < n;
        uint64_t key = 0;
        int val = 0, d = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in[i];
            d = sort_digit(mode, key, val, shift, T);
        }
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1;
            const uint64_t bb = ballot64(set);
            peers &= set ? bb : ~bb;
        }
        const int rank = __popcll(peers & lanemask_lt());
        if (valid && rank == 0) s_wc[w][d] = __popcll(peers);
        __syncthreads();
        int total;
        {
            int c0 = s_wc[0][t], c1 = s_wc[1][t], c2 = s_wc[2][t], c3 = s_wc[3][t];
            s_wc[0][t] = 0;
            s_wc[1][t] = c0;
            s_wc[2][t] = c0 + c1;
            s_wc[3][t] = c0 + c1 + c2;
            total = c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (valid) {
            const int pos = s_off[d] + s_wc[w][d] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        s_off[t] += total;
        __syncthreads();
    }
}

// Sorted bodies {x,y,z,m} for the walk leaves.
__global__ __launch_bounds__(kBlock) void k_gather(const int *__restrict__ perm,
                                                   const float *__restrict__ x,
                                                   const float *__restrict__ y,
                                                   const float *__restrict__ z,
                                                   const float *__restrict__ m, float4 *bodies,
                                                   int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int i = perm[s];
    bodies[s] = make_float4(x[i], y[i], z[i], m[i]);
}

// ===========================================================================
// Binary radix tree over the sorted keys of each disk (longest common prefix
// search, ties broken by position).  Internal node j of disk s covers a
// range of sorted positions; children are encoded >= 0 (internal, global
// index) or < 0 (leaf, ~global sorted position).
// ===========================================================================

__device__ inline int lcp(const uint64_t *__restrict__ k, int n, int a, int b)
{
    if (b < 0 || b >= n) return -1;
    uint64_t ka = k[a], kb = k[b];
    if (ka == kb) return 64 + __clz((unsigned)(a ^ b));
    return __clzll(ka ^ kb);
}

__global__ __launch_bounds__(kBlock) void k_karras(SegTable T, const uint64_t *__restrict__ keys,
                                                   int *left, int *right, int *par_int,
                                                   int *par_leaf, int *plen)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= T.ioff[T.nseg]) return;
    const int s = seg_of_internal(T, j);
    const int n = T.off[s + 1] - T.off[s];
    const uint64_t *k = keys + T.off[s];
    const int i = j - T.ioff[s];
    const int d = (lcp(k, n, i, i + 1) - lcp(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = lcp(k, n, i, i - d);
    int lmax = 2;
    while (lcp(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lcp(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int jj = i + l * d;
    const int dnode = lcp(k, n, i, jj);
    int sp = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (lcp(k, n, i, i + (sp + t) * d) > dnode) sp += t;
    } while (t > 1);
    const int gamma = i + sp * d + min(d, 0);
    const int lo = min(i, jj), hi = max(i, jj);
    int lc, rc;
    if (lo == gamma) {
        lc = ~(T.off[s] + gamma);
        par_leaf[T.off[s] + gamma] = j;
    } else {
        lc = T.ioff[s] + gamma;
        par_int[lc] = j;
    }
    if (hi == gamma + 1) {
        rc = ~(T.off[s] + gamma + 1);
        par_leaf[T.off[s] + gamma + 1] = j;
    } else {
        rc = T.ioff[s] + gamma + 1;
        par_int[rc] = j;
    }
    left[j] = lc;
    right[j] = rc;
    plen[j] = dnode;
    if (i == 0) par_int[j] = -1;
}

// Bottom-up mass / COM: each leaf climbs; the second thread to arrive at a
// node combines its two children and continues.
__global__ __launch_bounds__(kBlock) void k_com(SegTable T, const float4 *__restrict__ bodies,
                                                const int *__restrict__ par_leaf,
                                                const int *__restrict__ par_int,
                                                const int *__restrict__ left,
                                                const int *__restrict__ right, int *flags,
                                                float4 *bin_com)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= T.n_total) return;
    const int s = seg_of_body(T, p);
    if (T.off[s + 1] - T.off[s] < 2) return;
    int node = par_leaf[p];
    while (node >= 0) {
        int old = __hip_atomic_fetch_add(&flags[node], 1, __ATOMIC_ACQ_REL,
                                         __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        const int lc = left[node], rc = right[node];
        float4 a, b;
        if (lc < 0) a = bodies[~lc];
        else a = bin_com[lc];
        if (rc < 0) b = bodies[~rc];
        else b = bin_com[rc];
        const float m = a.w + b.w;
        float4 c;
        if (m > 0.f) {
            const float inv = 1.0f / m;
            c = make_float4((a.w * a.x + b.w * b.x) * inv, (a.w * a.y + b.w * b.y) * inv,
                            (a.w * a.z + b.w * b.z) * inv, m);
        } else {
            c = make_float4(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.f);
        }
        bin_com[node] = c;
        node = par_int[node];
    }
}

// Octree level of binary node j: floor(prefix bits / 3); identical keys
// (prefix >= 64) form level-21 cells.
__device__ inline int oct_level(int plen)
{
    return plen >= 64 ? kLevels : (plen - 1) / 3;
}

// A binary node is folded into its parent when both lie in the same digit.
__device__ inline bool is_oct(const int *__restrict__ par_int, const int *__restrict__ plen,
                              int j)
{
    const int p = par_int[j];
    if (p < 0) return true;
    const int lj = oct_level(plen[j]);
    return !(lj == oct_level(plen[p]) && lj < kLevels);
}

// Enumerate the octree children of octree node j in Morton order: DFS over
// the folded binary descendants (at most 7 for a digit, bucket nodes are
// binary).  f(child_code) is called for each child in order.
template <class F>
__device__ inline int oct_children(const int *__restrict__ left, const int *__restrict__ right,
                                   const int *__restrict__ par_int,
                                   const int *__restrict__ plen, int j, F f)
{
    int stack[16];
    int sp = 0, n = 0;
    stack[sp++] = right[j];
    stack[sp++] = left[j];
    while (sp > 0) {
        const int c = stack[--sp];
        if (c < 0 || is_oct(par_int, plen, c)) {
            f(n, c);
            ++n;
        } else {
            if (sp + 2 > 16) return -1;
            stack[sp++] = right[c];
            stack[sp++] = left[c];
        }
    }
    return n;
}

__global__ __launch_bounds__(kBlock) void k_oct_count(SegTable T, const int *__restrict__ left,
                                                      const int *__restrict__ right,
                                                      const int *__restrict__ par_int,
                                                      const int *__restrict__ plen, int *cnt,
                                                      int *first, int *alloc, int *maxlev,
                                                      int *err)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= T.ioff[T.nseg]) return;
    if (!is_oct(par_int, plen, j)) {
        cnt[j] = 0;
        return;
    }
    const int s = seg_of_internal(T, j);
    const int c = oct_children(left, right, par_int, plen, j, [](int, int) {});
    if (c < 0) {
        atomicOr(err, kErrTree);
        cnt[j] = 0;
        return;
    }
    cnt[j] = c;
    first[j] = T.noff[s] + 1 + atomicAdd(&alloc[s], c);
    atomicMax(&maxlev[s], oct_level(plen[j]));
}

// Walk records: children of octree node j into its block, and the root
// record of each disk at noff[s].
__global__ __launch_bounds__(kBlock) void k_oct_fill(SegTable T, const Root *__restrict__ roots,
                                                     const int *__restrict__ left,
                                                     const int *__restrict__ right,
                                                     const int *__restrict__ par_int,
                                                     const int *__restrict__ plen,
                                                     const int *__restrict__ cnt,
                                                     const int *__restrict__ first,
                                                     const float4 *__restrict__ bin_com,
                                                     const float4 *__restrict__ bodies,
                                                     float4 *rec_a, int4 *rec_b)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= T.ioff[T.nseg]) return;
    const int c_j = cnt[j];
    if (c_j == 0) return;
    const int s = seg_of_internal(T, j);
    const float rhs = roots[s].c[3];
    auto rec_of = [&](int code, float4 &a, int4 &b) {
        if (code < 0) {
            const int p = ~code;
            a = bodies[p];
            b = make_int4(__float_as_int(0.f), __float_as_int(-1.f), p, 0);
        } else {
            const float h = ldexpf(rhs, -oct_level(plen[code]));
            a = bin_com[code];
            b = make_int4(__float_as_int(h), __float_as_int(h * h), first[code], cnt[code]);
        }
    };
    if (par_int[j] < 0) {
        float4 a;
        int4 b;
        rec_of(j, a, b);
        rec_a[T.noff[s]] = a;
        rec_b[T.noff[s]] = b;
    }
    const int f0 = first[j];
    oct_children(left, right, par_int, plen, j, [&](int k, int code) {
        float4 a;
        int4 b;
        rec_of(code, a, b);
        rec_a[f0 + k] = a;
        rec_b[f0 + k] = b;
    });
}

// Disks of one body: the root record is the body itself.
__global__ void k_single_roots(SegTable T, const float4 *__restrict__ bodies, float4 *rec_a,
                               int4 *rec_b)
{
    const int s = threadIdx.x;
    if (s >= T.nseg) return;
    if (T.off[s + 1] - T.off[s] != 1) return;
    const int p = T.off[s];
    rec_a[T.noff[s]] = bodies[p];
    rec_b[T.noff[s]] = make_int4(__float_as_int(0.f), __float_as_int(-1.f), p, 0);
}

// ===========================================================================
// Launchers
// ===========================================================================

#define NBX_GRID(n) dim3(cdiv((n), kBlock)), dim3(kBlock), 0, st

void launch_half_kick(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                      const float *az, float dt_half, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_half_kick, NBX_GRID(n), vx, vy, vz, ax, ay, az, dt_half, n);
}
void launch_drift(float *x, float *y, float *z, const float *vx, const float *vy,
                  const float *vz, float dt, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_drift, NBX_GRID(n), x, y, z, vx, vy, vz, dt, n);
}
void launch_reset_acc(float *ax, float *ay, float *az, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_acc, NBX_GRID(n), ax, ay, az, n);
}
void launch_gas_drag(const float *x, const float *z, float *vx, float *vy, float *vz,
                     float decay, float star_mass, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_gas_drag, NBX_GRID(n), x, z, vx, vy, vz, decay, star_mass, n);
}
void launch_perturber(float *ax, float *ay, float *az, const float *x, const float *y,
                      const float *z, float px, float py, float pz, float pmass, float soft,
                      int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_perturber, NBX_GRID(n), ax, ay, az, x, y, z, px, py, pz, pmass, soft,
                       n);
}
void launch_fill_int(int *p, int v, int n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_int, NBX_GRID(n), p, v, n);
}
void launch_fill_float(float *p, float v, int n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_float, NBX_GRID(n), p, v, n);
}

void launch_kick_drift_bbox(const SegTable &T, float *x, float *y, float *z, float *vx,
                            float *vy, float *vz, const float *ax, const float *ay,
                            const float *az, float dt_half, float dt, Bufs &B, hipStream_t st)
{
    hipLaunchKernelGGL(k_kick_drift_bbox, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z,
                       vx, vy, vz, ax, ay, az, dt_half, dt, B.partial);
}

void launch_kick_drag(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                      const float *az, const float *dv, const float *x, const float *z,
                      float dt_half, int has_drag, float decay, float star_mass, int n,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_kick_drag, NBX_GRID(n), vx, vy, vz, ax, ay, az, dv, x, z, dt_half,
                       has_drag, decay, star_mass, n);
}

void launch_bbox(const SegTable &T, const float *x, const float *y, const float *z, Bufs &B,
                 hipStream_t st)
{
    hipLaunchKernelGGL(k_bbox, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z, B.partial);
}

void launch_root(const SegTable &T, Bufs &B, hipStream_t st)
{
    hipLaunchKernelGGL(k_root, dim3(T.nseg), dim3(kBlock), 0, st, T, B.partial, B.roots);
}

static void sort_pass(const SegTable &T, Bufs &B, int &cur, int shift, int mode, hipStream_t st)
{
    const int n = B.n, nt = B.n_tiles;
    hipLaunchKernelGGL(k_sort_hist, dim3(nt), dim3(kBlock), 0, st, B.keys[cur], B.vals[cur], n,
                       shift, mode, T, nt, B.hist);
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, st, B.hist, 256 * nt);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nt), dim3(kBlock), 0, st, B.keys[cur], B.vals[cur],
                       B.keys[cur ^ 1], B.vals[cur ^ 1], n, shift, mode, T, nt, B.hist);
    cur ^= 1;
}

void launch_build(const SegTable &T, const float *x, const float *y, const float *z,
                  const float *mass, Bufs &B, hipStream_t st)
{
    const int n = T.n_total;
    hipLaunchKernelGGL(k_keys, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z, B.roots,
                       B.keys[0], B.vals[0]);
    int cur = 0;
    for (int shift = 0; shift < 64; shift += 8) sort_pass(T, B, cur, shift, 0, st);
    if (T.nseg > 1) {
        sort_pass(T, B, cur, 0, 1, st);
        // keep the sorted arrays in buffer 0
        if (cur != 0) {
            (void)hipMemcpyAsync(B.keys[0], B.keys[1], sizeof(uint64_t) * n,
                                 hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(B.vals[0], B.vals[1], sizeof(int) * n, hipMemcpyDeviceToDevice,
                                 st);
        }
    }
    hipLaunchKernelGGL(k_gather, NBX_GRID(n), B.vals[0], x, y, z, mass, B.bodies, n);
    const int ni = T.ioff[T.nseg];
    if (ni > 0)
        hipLaunchKernelGGL(k_karras, NBX_GRID(ni), T, B.keys[0], B.left, B.right, B.par_int,
                           B.par_leaf, B.plen);
}

void launch_summarize(const SegTable &T, Bufs &B, hipStream_t st)
{
    const int n = T.n_total, ni = T.ioff[T.nseg];
    launch_fill_int(B.flags, 0, ni, st);
    launch_fill_int(B.alloc, 0, kMaxSeg, st);
    launch_fill_int(B.maxlev, 0, kMaxSeg, st);
    if (ni > 0) {
        hipLaunchKernelGGL(k_com, NBX_GRID(n), T, B.bodies, B.par_leaf, B.par_int, B.left,
                           B.right, B.flags, B.bin_com);
        hipLaunchKernelGGL(k_oct_count, NBX_GRID(ni), T, B.left, B.right, B.par_int, B.plen,
                           B.cnt, B.first, B.alloc, B.maxlev, B.err);
        hipLaunchKernelGGL(k_oct_fill, NBX_GRID(ni), T, B.roots, B.left, B.right, B.par_int,
                           B.plen, B.cnt, B.first, B.bin_com, B.bodies, B.rec_a, B.rec_b);
    }
    hipLaunchKernelGGL(k_single_roots, dim3(1), dim3(64), 0, st, T, B.bodies, B.rec_a, B.rec_b);
}

}  // namespace nbx
