// nbx_adjoint.hip -- reverse mode of the collision-free disk step on gfx950.
//
// Forward step n (main.cu:101-140 with collisions off):
//   v_h     = v_n + a_n dt/2,          x_{n+1} = x_n + v_h dt
//   a_{n+1} = A(x_{n+1}, t_{n+1})      (tree walk + star + perturber)
//   v'      = v_h + a_{n+1} dt/2,      v_{n+1} = Drag(x_{n+1}, v')
// Reverse step n, given the adjoints xb, vb of x_{n+1}, v_{n+1} and ab of
// a_{n+1} (its share from the first kick of step n+1):
//   k_adj_kick_drag : vb' = (dDrag/dv')^T vb,  xb += (dDrag/dx)^T vb,
//                     ab += vb' dt/2,  vb = vb'
//   adjoint force   : xb += (dA/dx)^T ab,  dL/d(m_p, r_p) += (dA/dp)^T ab
//   k_adj_drift_kick: vb += xb dt,  ab = vb dt/2     (adjoints of x_n, v_n, a_n)
//
// The force adjoint freezes the interaction lists of the forward pass: the
// wave walk takes the forward walk's decisions on the same positions (same
// records, same fixed-rounding opening test).  For a target i and an
// accepted record c with d = X_c - x_i and s = |d|^2 + eps^2, the pull
// m_c d / s^{3/2} has the symmetric Jacobian J(d) = I/s^{3/2} - 3 d d^T/s^{5/2}:
//   target side: xb_i -= m_c J(d) ab_i
//   record side: G_c  += J(d) ab_i, and X_c = sum_{j in c} m_j x_j / M_c
//                gives xb_j += m_j sum_{c containing j} G_c.
// A record covers a contiguous range of sorted positions, so the last sum
// is a range add: differences at the range ends, then one prefix scan.
// Arithmetic per interaction is FP32; every sum is accumulated in FP64.
#include "nbx_wave.hpp"

namespace nbx {

// J(d) u for the softened pull d / s^{3/2}.
__device__ inline void jmul(float dx, float dy, float dz, float s, float ux, float uy, float uz,
                            float &ox, float &oy, float &oz)
{
    const float rinv = 1.0f / sqrtf(s);
    const float r3 = rinv * rinv * rinv;
    const float r5 = r3 * rinv * rinv;
    const float du = 3.0f * (dx * ux + dy * uy + dz * uz) * r5;
    ox = ux * r3 - dx * du;
    oy = uy * r3 - dy * du;
    oz = uz * r3 - dz * du;
}

template <int TG>
__device__ inline double slot_sum_d(double v)
{
#pragma unroll
    for (int o = TG; o < kWave; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int TG>
struct AdjVisitor {
    static constexpr bool kRecId = true;
    float bx, by, bz, ith2, eps2;  // ith2 = theta^-2 (the walk stages h^2 ith2)
    int self;
    float ux, uy, uz;     // abar of the target
    double gx, gy, gz;    // target side: -sum m_c J(d) abar
    double *G;            // [3 n_rec] record side: sum J(d) abar
    __device__ inline bool visit_masked(const Rec &r, bool mine)
    {
        const float4 &a = r.a;
        const int4 &b = r.b;
        const bool leaf = b.w == 0;
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float r2 = dist2(dx, dy, dz);
        const bool open = mine && !leaf && !(__int_as_float(b.y) < r2);
        const bool take = mine && !open && !(leaf && b.z == self);
        float jx, jy, jz;
        jmul(dx, dy, dz, r2 + eps2, ux, uy, uz, jx, jy, jz);
        if (!take) jx = jy = jz = 0.f;
        gx -= (double)(a.w * jx);
        gy -= (double)(a.w * jy);
        gz -= (double)(a.w * jz);
        // sum over the TG targets of this record slot, one FP64 atomic per record
        double sx = jx, sy = jy, sz = jz;
#pragma unroll
        for (int o = 1; o < TG; o <<= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
            sz += __shfl_xor(sz, o, 64);
        }
        const uint64_t tk = ballot64(take);
        const int lane = __lane_id();
        const int slot0 = lane - lane % TG;
        const uint64_t smask = TG == 64 ? ~0ull : ((1ull << TG) - 1ull);
        if (lane == slot0 && ((tk >> slot0) & smask)) {
            atomicAdd(G + 3 * r.id + 0, sx);
            atomicAdd(G + 3 * r.id + 1, sy);
            atomicAdd(G + 3 * r.id + 2, sz);
        }
        return !open;  // the walk opens the record where mine and not returned
    }
};

// Perturber of one force evaluation: position and mass as in the forward
// walk, and dX_p/dr_p of its orbit at that time.
struct AdjPert {
    float px, py, pz, pmass;
    float drx, dry, drz;
};

template <int TG>
__global__ __launch_bounds__(kWalkBlock) void k_adj_walk(
    SegTable T, WaveMap M, AdjPert P, const float4 *__restrict__ ra, const int4 *__restrict__ rb,
    const float4 *__restrict__ bodies, const int *__restrict__ perm, float theta2, float eps2,
    float softening, float star_mass, const double *__restrict__ ab, double *xb, double *G,
    double *pg, int *err)
{
    __shared__ WaveLdsIdT<(TG > 32)> lds[kWalkWaves];
    const int wib = threadIdx.x >> 6;
    const int w = blockIdx.x * kWalkWaves + wib;
    if (w >= M.wave[T.nseg]) return;  // uniform per wave
    const int s = seg_of_wmap(T, M, w);
    const int n = T.n_total;
    const int lane = __lane_id(), tt = lane % TG, rr = lane / TG;
    const int pos = T.off[s] + (w - M.wave[s]) * TG + tt;
    const bool valid = pos < T.off[s + 1];
    const float4 b = bodies[valid ? pos : T.off[s]];
    const int i = perm[valid ? pos : T.off[s]];
    const float ux = valid ? (float)ab[i] : 0.f;
    const float uy = valid ? (float)ab[n + i] : 0.f;
    const float uz = valid ? (float)ab[2 * n + i] : 0.f;
    AdjVisitor<TG> v{b.x, b.y, b.z, 1.0f / theta2, eps2, pos, ux, uy, uz, 0.0, 0.0, 0.0, G};
    const uint64_t active = ballot64(valid && rr == 0);
    wave_walk<TG, false>(ra, rb, nullptr, T.noff[s], active, lds[wib], err, v);
    double gx = slot_sum_d<TG>(v.gx), gy = slot_sum_d<TG>(v.gy), gz = slot_sum_d<TG>(v.gz);
    double pm = 0.0, pr = 0.0;
    if (valid && rr == 0) {
        float jx, jy, jz;
        // star at the origin, pull M d / s^{3/2} with d = -x (BHTree.cu:547-553)
        const float ss = dist2(b.x, b.y, b.z) + softening * softening;
        jmul(-b.x, -b.y, -b.z, ss, ux, uy, uz, jx, jy, jz);
        gx -= (double)(star_mass * jx);
        gy -= (double)(star_mass * jy);
        gz -= (double)(star_mass * jz);
        if (P.pmass > 0.f) {  // Perturber.cu:131-143
            const float dx = P.px - b.x, dy = P.py - b.y, dz = P.pz - b.z;
            const float sp = dist2(dx, dy, dz) + softening * softening;
            jmul(dx, dy, dz, sp, ux, uy, uz, jx, jy, jz);
            gx -= (double)(P.pmass * jx);
            gy -= (double)(P.pmass * jy);
            gz -= (double)(P.pmass * jz);
            const float k3 = 1.0f / (sp * sqrtf(sp));
            pm = (double)((dx * ux + dy * uy + dz * uz) * k3);
            pr = (double)(P.pmass * (jx * P.drx + jy * P.dry + jz * P.drz));
        }
        xb[i] += gx;
        xb[n + i] += gy;
        xb[2 * n + i] += gz;
    }
    pm = wave_sum_d(pm);
    pr = wave_sum_d(pr);
    if (lane == 0 && (pm != 0.0 || pr != 0.0)) {
        atomicAdd(pg + 0, pm);
        atomicAdd(pg + 1, pr);
    }
}

// Range adds of the record sums: D[lo] += G_c, D[hi + 1] -= G_c.
__global__ void k_adj_nodes(const double *__restrict__ G, const int2 *__restrict__ rr, double *D,
                            int n_rec)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rec) return;
    const double gx = G[3 * r], gy = G[3 * r + 1], gz = G[3 * r + 2];
    if (gx == 0.0 && gy == 0.0 && gz == 0.0) return;
    const int2 q = rr[r];
    atomicAdd(D + 3 * q.x + 0, gx);
    atomicAdd(D + 3 * q.x + 1, gy);
    atomicAdd(D + 3 * q.x + 2, gz);
    atomicAdd(D + 3 * (q.y + 1) + 0, -gx);
    atomicAdd(D + 3 * (q.y + 1) + 1, -gy);
    atomicAdd(D + 3 * (q.y + 1) + 2, -gz);
}

constexpr int kAdjScanItems = 4;
constexpr int kAdjScanTile = kBlock * kAdjScanItems;

// Scan of 3-vectors of doubles inside one block: v becomes the exclusive
// prefix of the per-thread values, tot the block total.
__device__ inline void block_scan3(double v[3], double tot[3])
{
    __shared__ double s[3][kBlock];
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 3; ++c) s[c][t] = v[c];
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        double u[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = t >= o ? s[c][t - o] : 0.0;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c][t] += u[c];
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        tot[c] = s[c][kBlock - 1];
        v[c] = t > 0 ? s[c][t - 1] : 0.0;  // exclusive prefix of this thread
    }
    __syncthreads();
}

// Pass 1: per tile, inclusive prefix of D in place and the tile total.
__global__ __launch_bounds__(kBlock) void k_adj_scan_tiles(double *D, double *blk, int n)
{
    const int base = blockIdx.x * kAdjScanTile + threadIdx.x * kAdjScanItems;
    double run[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kAdjScanItems; ++k)
        if (base + k < n)
#pragma unroll
            for (int c = 0; c < 3; ++c) run[c] += D[3 * (base + k) + c];
    double tot[3];
    block_scan3(run, tot);
#pragma unroll
    for (int k = 0; k < kAdjScanItems; ++k)
        if (base + k < n)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                run[c] += D[3 * (base + k) + c];
                D[3 * (base + k) + c] = run[c];
            }
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) blk[3 * blockIdx.x + c] = tot[c];
}

// Pass 2 (one block): exclusive prefix of the tile totals, in place.
__global__ __launch_bounds__(kBlock) void k_adj_scan_totals(double *blk, int nb)
{
    double carry[3] = {0.0, 0.0, 0.0};
    for (int b0 = 0; b0 < nb; b0 += kBlock) {
        const int b = b0 + threadIdx.x;
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = b < nb ? blk[3 * b + c] : 0.0;
        double tot[3];
        block_scan3(v, tot);
        if (b < nb)
#pragma unroll
            for (int c = 0; c < 3; ++c) blk[3 * b + c] = carry[c] + v[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) carry[c] += tot[c];
    }
}

// Pass 3: xb[perm[j]] += m_j (inclusive prefix of D at j).
__global__ __launch_bounds__(kBlock) void k_adj_scan_apply(const double *__restrict__ D,
                                                           const double *__restrict__ blk,
                                                           const float4 *__restrict__ bodies,
                                                           const int *__restrict__ perm,
                                                           double *xb, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int tile = j / kAdjScanTile;
    const double m = (double)bodies[j].w;
    const int i = perm[j];
#pragma unroll
    for (int c = 0; c < 3; ++c) xb[c * n + i] += m * (D[3 * j + c] + blk[3 * tile + c]);
}

__global__ void k_fill_d(double *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Reverse of the second kick and the drag (Integrator.cu:40-76): the drag
// maps v' to decay v' + (1 - decay) sign(v_phi) sqrt(M) r^{-3/2} (-z, 0, x)
// in the XZ plane about the star, so dDrag/dv' = decay I and the position
// derivative comes from the Keplerian target velocity.  v' is recomputed
// with the forward's float expressions (its sign of v_phi decides the branch).
__global__ void k_adj_kick_drag(const float *__restrict__ x1, const float *__restrict__ z1,
                                const float *__restrict__ vx0, const float *__restrict__ vy0,
                                const float *__restrict__ vz0, const float *__restrict__ ax0,
                                const float *__restrict__ ay0, const float *__restrict__ az0,
                                const float *__restrict__ ax1, const float *__restrict__ ay1,
                                const float *__restrict__ az1, float dt_half, int has_drag,
                                float decay, float star_mass, double *xb, double *vb, double *ab,
                                int n)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double bx = vb[i], by = vb[n + i], bz = vb[2 * n + i];
    if (has_drag) {
        const float xi = x1[i], zi = z1[i];
        const float r_xz2 = xi * xi + zi * zi;
        if (!(r_xz2 < 1e-12f)) {
            const float vpx = (vx0[i] + ax0[i] * dt_half) + ax1[i] * dt_half;
            const float vpz = (vz0[i] + az0[i] * dt_half) + az1[i] * dt_half;
            const float inv_r = 1.0f / sqrtf(r_xz2);
            const float v_phi = (vpx * (-zi) + vpz * xi) * inv_r;
            const double sg = v_phi >= 0.0f ? 1.0 : -1.0;
            const double xd = xi, zd = zi;
            const double r = sqrt(xd * xd + zd * zd);
            const double rm32 = 1.0 / (r * sqrt(r));
            const double rm72 = rm32 / (r * r);
            const double c = sg * (1.0 - (double)decay) * sqrt((double)star_mass);
            // g = (-z, x) r^{-3/2}: partial derivatives in x and z
            const double gxx = 1.5 * zd * xd * rm72, gxz = -rm32 + 1.5 * zd * zd * rm72;
            const double gzx = rm32 - 1.5 * xd * xd * rm72, gzz = -1.5 * xd * zd * rm72;
            xb[i] += c * (bx * gxx + bz * gzx);
            xb[2 * n + i] += c * (bx * gxz + bz * gzz);
            bx *= (double)decay;
            by *= (double)decay;
            bz *= (double)decay;
        }
    }
    vb[i] = bx;
    vb[n + i] = by;
    vb[2 * n + i] = bz;
    ab[i] += bx * (double)dt_half;
    ab[n + i] += by * (double)dt_half;
    ab[2 * n + i] += bz * (double)dt_half;
}

// Reverse of the drift and the first kick.
__global__ void k_adj_drift_kick(const double *__restrict__ xb, double *vb, double *ab, float dt,
                                 float dt_half, int n3)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const double v = vb[i] + xb[i] * (double)dt;
    vb[i] = v;
    ab[i] = v * (double)dt_half;
}

// ===========================================================================
// Launchers
// ===========================================================================

#define NBX_GRID(n) dim3(cdiv((n), kBlock)), dim3(kBlock), 0, st

template <int TG>
static void launch_adj_walk_tg(const SegTable &T, const AdjWalkArgs &A, Bufs &B, hipStream_t st)
{
    WaveMap M;
    M.wave[0] = 0;
    for (int s = 0; s < kMaxSeg; ++s)
        M.wave[s + 1] = M.wave[s] + (s < T.nseg ? cdiv(T.off[s + 1] - T.off[s], TG) : 0);
    const int waves = M.wave[T.nseg];
    const AdjPert P{A.px, A.py, A.pz, A.pmass, A.drx, A.dry, A.drz};
    hipLaunchKernelGGL(k_adj_walk<TG>, dim3(cdiv(waves, kWalkWaves)), dim3(kWalkBlock), 0, st, T,
                       M, P, B.rec_a, B.rec_b, B.bodies, B.vals[B.cur], A.theta2, A.eps2,
                       A.softening, A.star_mass, A.ab, A.xb, B.adj_g, B.adj_p, B.err);
}

void launch_adj_force(const SegTable &T, const AdjWalkArgs &A, Bufs &B, hipStream_t st)
{
    const int n = T.n_total;
    hipLaunchKernelGGL(k_fill_d, NBX_GRID(3 * B.n_rec), B.adj_g, 3 * B.n_rec);
    hipLaunchKernelGGL(k_fill_d, NBX_GRID(3 * (n + 1)), B.adj_d, 3 * (n + 1));
    hipLaunchKernelGGL(k_fill_d, NBX_GRID(4), B.adj_p, 4);
    switch (A.targets) {
    case 8: launch_adj_walk_tg<8>(T, A, B, st); break;
    case 16: launch_adj_walk_tg<16>(T, A, B, st); break;
    case 32: launch_adj_walk_tg<32>(T, A, B, st); break;
    default: launch_adj_walk_tg<64>(T, A, B, st); break;
    }
    hipLaunchKernelGGL(k_adj_nodes, NBX_GRID(B.n_rec), B.adj_g, B.rec_r, B.adj_d, B.n_rec);
    const int nb = cdiv(n, kAdjScanTile);
    hipLaunchKernelGGL(k_adj_scan_tiles, dim3(nb), dim3(kBlock), 0, st, B.adj_d, B.adj_blk, n);
    hipLaunchKernelGGL(k_adj_scan_totals, dim3(1), dim3(kBlock), 0, st, B.adj_blk, nb);
    hipLaunchKernelGGL(k_adj_scan_apply, NBX_GRID(n), B.adj_d, B.adj_blk, B.bodies,
                       B.vals[B.cur], A.xb, n);
}

int adj_scan_blocks(int n) { return cdiv(n, kAdjScanTile); }

void launch_adj_kick_drag(const float *x1, const float *z1, const float *vx0, const float *vy0,
                          const float *vz0, const float *ax0, const float *ay0, const float *az0,
                          const float *ax1, const float *ay1, const float *az1, float dt_half,
                          int has_drag, float decay, float star_mass, double *xb, double *vb,
                          double *ab, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_kick_drag, NBX_GRID(n), x1, z1, vx0, vy0, vz0, ax0, ay0, az0, ax1,
                       ay1, az1, dt_half, has_drag, decay, star_mass, xb, vb, ab, n);
}

void launch_adj_drift_kick(const double *xb, double *vb, double *ab, float dt, float dt_half,
                           int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_drift_kick, NBX_GRID(3 * n), xb, vb, ab, dt, dt_half, 3 * n);
}

}  // namespace nbx
