// nbx_tree.hip -- octree construction and integrator kernels for gfx950.
//
// Build pipeline (one launch per phase for all disks of a batch):
//   kick-drift with per-block boxes -> key pass: root cube, 63-bit keys, each
//   body scattered to its sort bucket (splitters from the last build's
//   quantiles) -> bucket sort in LDS (sorted keys, indices and bodies) ->
//   FP64 prefix scan of {m, m x, m y, m z} -> binary radix tree (one thread
//   per internal node) with node mass / COM from the prefix, octree flags
//   and cell geometry -> child counts -> records.  Disks above 2^20
//   bodies (or a set digit width) sort with LSD radix passes instead.
// Every phase is HBM-streaming integer/float work; nothing here is a GEMM.
#include <float.h>

#include <algorithm>

#include "nbx_device.hpp"
#include "nbx_launch.hpp"

namespace nbx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ===========================================================================
// Integrator kernels.  Float expressions are written exactly as in
// Integrator.cu / Perturber.cu and compiled without contraction, so every
// operation rounds as the reference source reads without fused
// multiply-adds (the oracle evaluates them the same way).  nvcc contracts
// a * b + c into an FMA by default, so a reference binary may round some of
// these expressions differently in the last bit.
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

// Fused first half-kick + drift (main.cu:104-107) with the bbox of the
// drifted positions (BHTree.cu:94-133) in the same pass over HBM.
__device__ void bbox_post(float lx, float ly, float lz, float hx, float hy, float hz,
                          float *partial);

__global__ __launch_bounds__(kBlock) void k_kick_drift_bbox(
    SegTable T, float *x, float *y, float *z, float *vx, float *vy, float *vz,
    const float *__restrict__ ax, const float *__restrict__ ay, const float *__restrict__ az,
    float dt_half, float dt, float *partial, LptDev L)
{
#pragma clang fp contract(off)
    lpt_count(L);
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
    bbox_post(lx, ly, lz, hx, hy, hz, partial);
}

// Collision delta, second half-kick and gas drag (main.cu:131-139 after the
// impulse pass): v += dv; v += a*dt_half; drag.
__global__ void k_kick_drag(float *vx, float *vy, float *vz, const float *__restrict__ ax,
                            const float *__restrict__ ay, const float *__restrict__ az,
                            const float *__restrict__ dv, const float *__restrict__ x,
                            const float *__restrict__ z, float dt_half, int has_drag,
                            float decay, float star_mass, int n, int *step_ctr)
{
#pragma clang fp contract(off)
    if (step_ctr && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(step_ctr, 1);
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

// The end of one step and the start of the next in one pass: v += dv;
// v += a dt_half; drag (main.cu:131-139), then v += a dt_half; x += v dt
// (main.cu:104-107, the same acceleration: no force pass lies between) and
// the per-block boxes of the new positions.  Every value rounds as in the
// two kernels it replaces.  step_ctr (optional) advances by one.
__global__ __launch_bounds__(kBlock) void k_drag_kick_drift(
    SegTable T, float *x, float *y, float *z, float *vx, float *vy, float *vz,
    const float *__restrict__ ax, const float *__restrict__ ay, const float *__restrict__ az,
    const float *__restrict__ dv, float dt_half, float dt, int has_drag, float decay,
    float star_mass, int *step_ctr, float *partial)
{
#pragma clang fp contract(off)
    if (step_ctr && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(step_ctr, 1);
    const int b = blockIdx.x;
    const int s = seg_of_block(T, b);
    const int i = T.off[s] + (b - T.blk[s]) * kBlock + threadIdx.x;
    const int n = T.n_total;
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
    float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    if (i < T.off[s + 1]) {
        const float aX = ax[i], aY = ay[i], aZ = az[i];
        float u = vx[i], v = vy[i], w = vz[i];
        float X = x[i], Y = y[i], Z = z[i];
        if (dv) {
            u += dv[i];
            v += dv[n + i];
            w += dv[2 * n + i];
        }
        if (dt_half != 0.f) {  // as k_kick_drag
            u += aX * dt_half;
            v += aY * dt_half;
            w += aZ * dt_half;
        }
        if (has_drag) gas_drag_one(X, Z, u, v, w, decay, star_mass);
        u = u + aX * dt_half;
        v = v + aY * dt_half;
        w = w + aZ * dt_half;
        X = X + u * dt;
        Y = Y + v * dt;
        Z = Z + w * dt;
        vx[i] = u; vy[i] = v; vz[i] = w;
        x[i] = X; y[i] = Y; z[i] = Z;
        lx = hx = X;
        ly = hy = Y;
        lz = hz = Z;
    }
    bbox_post(lx, ly, lz, hx, hy, hz, partial);
}

// ---------------------------------------------------------------------------
// Compensated mode (no reference counterpart): the same KDK / drag sequence
// evaluated in FP64 on X = hi + lo, storing hi = round-to-float(X) in the
// caller's state and lo = X - hi in the tree.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_kick_drift_bbox_c(
    SegTable T, float *x, float *y, float *z, float *vx, float *vy, float *vz,
    const float *__restrict__ ax, const float *__restrict__ ay, const float *__restrict__ az,
    float dt_half, float dt, float *lo, float *partial, LptDev L)
{
    lpt_count(L);
    const int b = blockIdx.x;
    const int s = seg_of_block(T, b);
    const int i = T.off[s] + (b - T.blk[s]) * kBlock + threadIdx.x;
    const int n = T.n_total;
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
    float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    if (i < T.off[s + 1]) {
        float *xs[3] = {x, y, z};
        float *vs[3] = {vx, vy, vz};
        const float *as[3] = {ax, ay, az};
        float o[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double V = (double)vs[k][i] + (double)lo[(3 + k) * n + i];
            V += (double)as[k][i] * (double)dt_half;
            double X = (double)xs[k][i] + (double)lo[k * n + i];
            X += V * (double)dt;
            const float xh = (float)X, vh = (float)V;
            xs[k][i] = xh;
            vs[k][i] = vh;
            lo[k * n + i] = (float)(X - (double)xh);
            lo[(3 + k) * n + i] = (float)(V - (double)vh);
            o[k] = xh;
        }
        lx = hx = o[0];
        ly = hy = o[1];
        lz = hz = o[2];
    }
    bbox_post(lx, ly, lz, hx, hy, hz, partial);
}

__device__ inline void gas_drag_one_d(double xi, double zi, double &vxi, double &vyi,
                                      double &vzi, double decay, double star_mass)
{
    const double r_xz2 = xi * xi + zi * zi;
    if (r_xz2 < 1e-12) return;
    const double inv_r = 1.0 / sqrt(r_xz2);
    const double v_r = (vxi * xi + vzi * zi) * inv_r;
    const double v_phi = (vxi * (-zi) + vzi * xi) * inv_r;
    const double v_kep_mag = sqrt(star_mass * inv_r);
    const double v_kep = (v_phi >= 0.0) ? v_kep_mag : -v_kep_mag;
    const double v_r_new = v_r * decay;
    const double v_phi_new = v_kep + (v_phi - v_kep) * decay;
    vxi = v_r_new * (xi * inv_r) + v_phi_new * (-zi * inv_r);
    vzi = v_r_new * (zi * inv_r) + v_phi_new * (xi * inv_r);
    vyi = vyi * decay;
}

__global__ void k_kick_drag_c(float *vx, float *vy, float *vz, const float *__restrict__ ax,
                              const float *__restrict__ ay, const float *__restrict__ az,
                              const float *__restrict__ dv, const float *__restrict__ x,
                              const float *__restrict__ z, float dt_half, int has_drag,
                              double decay, float star_mass, int n, int *step_ctr, float *lo)
{
    if (step_ctr && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(step_ctr, 1);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double V[3] = {(double)vx[i] + lo[3 * n + i], (double)vy[i] + lo[4 * n + i],
                   (double)vz[i] + lo[5 * n + i]};
    const float *as[3] = {ax, ay, az};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (dv) V[k] += (double)dv[k * n + i];
        V[k] += (double)as[k][i] * (double)dt_half;
    }
    if (has_drag)
        gas_drag_one_d((double)x[i] + lo[i], (double)z[i] + lo[2 * n + i], V[0], V[1], V[2],
                       decay, (double)star_mass);
    float *vs[3] = {vx, vy, vz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float h = (float)V[k];
        vs[k][i] = h;
        lo[(3 + k) * n + i] = (float)(V[k] - (double)h);
    }
}

// ---------------------------------------------------------------------------
// FP64 step (no reference counterpart; the reference of the precision
// study): the state lives in s64 = [x | y | z | vx | vy | vz | ax | ay | az]
// (9 n doubles, body order) and every update is an FP64 operation; the
// caller's float arrays receive the rounded state after each pass.
// ---------------------------------------------------------------------------
__global__ void k_load64(const float *__restrict__ x, const float *__restrict__ y,
                         const float *__restrict__ z, const float *__restrict__ vx,
                         const float *__restrict__ vy, const float *__restrict__ vz, double *s64,
                         int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *src[6] = {x, y, z, vx, vy, vz};
#pragma unroll
    for (int k = 0; k < 6; ++k) s64[(size_t)k * n + i] = (double)src[k][i];
}

// The state of every body as doubles, [x | y | z | vx | vy | vz | ax | ay |
// az]: the FP64 state (s64), hi + lo (lo; accelerations are float), or the
// float state.
__global__ void k_state64(const float *__restrict__ x, const float *__restrict__ y,
                          const float *__restrict__ z, const float *__restrict__ vx,
                          const float *__restrict__ vy, const float *__restrict__ vz,
                          const float *__restrict__ ax, const float *__restrict__ ay,
                          const float *__restrict__ az, const float *__restrict__ lo,
                          const double *__restrict__ s64, int n, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *src[9] = {x, y, z, vx, vy, vz, ax, ay, az};
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const size_t q = (size_t)k * n + i;
        out[q] = s64 ? s64[q] : (double)src[k][i] + (lo && k < 6 ? (double)lo[q] : 0.0);
    }
}

__global__ __launch_bounds__(kBlock) void k_kick_drift_bbox_d(SegTable T, float *x, float *y,
                                                              float *z, float *vx, float *vy,
                                                              float *vz, double dt_half,
                                                              double dt, double *s64,
                                                              float *partial)
{
    const int b = blockIdx.x;
    const int s = seg_of_block(T, b);
    const int i = T.off[s] + (b - T.blk[s]) * kBlock + threadIdx.x;
    const size_t n = T.n_total;
    float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
    float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
    if (i < T.off[s + 1]) {
        float *xs[3] = {x, y, z};
        float *vs[3] = {vx, vy, vz};
        float o[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double V = s64[(3 + k) * n + i] + s64[(6 + k) * n + i] * dt_half;
            const double X = s64[k * n + i] + V * dt;
            s64[(3 + k) * n + i] = V;
            s64[k * n + i] = X;
            vs[k][i] = (float)V;
            xs[k][i] = o[k] = (float)X;
        }
        lx = hx = o[0];
        ly = hy = o[1];
        lz = hz = o[2];
    }
    bbox_post(lx, ly, lz, hx, hy, hz, partial);
}

__global__ void k_kick_drag_d(float *vx, float *vy, float *vz, double dt_half, int has_drag,
                              double decay, double star_mass, int n, int *step_ctr, double *s64)
{
    if (step_ctr && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(step_ctr, 1);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t m = n;
    double V[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) V[k] = s64[(3 + k) * m + i] + s64[(6 + k) * m + i] * dt_half;
    if (has_drag) gas_drag_one_d(s64[i], s64[2 * m + i], V[0], V[1], V[2], decay, star_mass);
    float *vs[3] = {vx, vy, vz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s64[(3 + k) * m + i] = V[k];
        vs[k][i] = (float)V[k];
    }
}

// FP64 diagnostics: block partial sums then one double atomic per block.
__global__ __launch_bounds__(kBlock) void k_diag(const float *__restrict__ x,
                                                 const float *__restrict__ y,
                                                 const float *__restrict__ z,
                                                 const float *__restrict__ vx,
                                                 const float *__restrict__ vy,
                                                 const float *__restrict__ vz,
                                                 const float *__restrict__ m,
                                                 const float *__restrict__ lo,
                                                 const double *__restrict__ s64, int n,
                                                 double star_mass, double soft, double *out)
{
    __shared__ double s[kBlock / kWave][6];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        double X = x[i], Y = y[i], Z = z[i], U = vx[i], V = vy[i], W = vz[i];
        if (s64) {
            X = s64[i]; Y = s64[n + i]; Z = s64[2 * n + i];
            U = s64[3 * n + i]; V = s64[4 * n + i]; W = s64[5 * n + i];
        } else if (lo) {
            X += lo[i]; Y += lo[n + i]; Z += lo[2 * n + i];
            U += lo[3 * n + i]; V += lo[4 * n + i]; W += lo[5 * n + i];
        }
        const double mm = m[i];
        e[0] = 0.5 * mm * (U * U + V * V + W * W);
        e[1] = -mm * star_mass / sqrt(X * X + Y * Y + Z * Z + soft * soft);
        e[2] = mm * (Y * W - Z * V);
        e[3] = mm * (Z * U - X * W);
        e[4] = mm * (X * V - Y * U);
        e[5] = mm;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) e[k] += __shfl_xor(e[k], o, 64);
    if (__lane_id() == 0)
        for (int k = 0; k < 6; ++k) s[threadIdx.x >> 6][k] = e[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = 0;
        for (int w = 0; w < kBlock / kWave; ++w) v += s[w][threadIdx.x];
        const int slot[6] = {0, 1, 2, 3, 4, 6};
        atomicAdd(&out[slot[threadIdx.x]], v);
    }
}

// Pairwise softened potential -1/2 sum_{i != j} m_i m_j / sqrt(r^2 + eps^2) in
// FP64, sources staged through LDS in tiles of kBlock.
__global__ __launch_bounds__(kBlock) void k_diag_self(const float *__restrict__ x,
                                                      const float *__restrict__ y,
                                                      const float *__restrict__ z,
                                                      const float *__restrict__ m,
                                                      const float *__restrict__ lo,
                                                      const double *__restrict__ s64, int n,
                                                      double soft, double *out)
{
    __shared__ double sx[kBlock], sy[kBlock], sz[kBlock], sm[kBlock];
    __shared__ double red[kBlock / kWave];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // position of body j: the FP64 state, hi + lo, or the float state
    auto pos = [&](int j, int k) {
        const float *f = k == 0 ? x : (k == 1 ? y : z);
        return s64 ? s64[(size_t)k * n + j] : (double)f[j] + (lo ? (double)lo[k * n + j] : 0.0);
    };
    double X = 0, Y = 0, Z = 0, M = 0;
    if (i < n) {
        X = pos(i, 0);
        Y = pos(i, 1);
        Z = pos(i, 2);
        M = m[i];
    }
    const double e2 = soft * soft;
    double phi = 0;
    for (int base = 0; base < n; base += kBlock) {
        const int j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = pos(j, 0);
            sy[threadIdx.x] = pos(j, 1);
            sz[threadIdx.x] = pos(j, 2);
            sm[threadIdx.x] = m[j];
        } else {
            sx[threadIdx.x] = sy[threadIdx.x] = sz[threadIdx.x] = 0;
            sm[threadIdx.x] = 0;
        }
        __syncthreads();
        const int lim = min(kBlock, n - base);
        for (int k = 0; k < lim; ++k) {
            if (base + k == i) continue;
            const double dx = sx[k] - X, dy = sy[k] - Y, dz = sz[k] - Z;
            phi += sm[k] / sqrt(dx * dx + dy * dy + dz * dz + e2);
        }
        __syncthreads();
    }
    double v = i < n ? -0.5 * M * phi : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / kWave; ++w) t += red[w];
        atomicAdd(&out[5], t);
    }
}

__global__ void k_fill_double(double *p, double v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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
// Bounding box and root cube.  Every block stores its min/max; the key pass
// reduces them per disk, pads the box by 1.001 about its midpoint
// (BHTree.cu:440-446) and forms the root cube (BHTree.cu:187-191).
// ===========================================================================

__device__ void root_from_box(const float *lo_hi, Root *out)
{
#pragma clang fp contract(off)
    float bmin[3], bmax[3];
    for (int k = 0; k < 3; ++k) {
        float lo = lo_hi[k], hi = lo_hi[3 + k];
        float mid = (lo + hi) * 0.5f;
        float half = (hi - lo) * 0.5f * 1.001f;
        bmin[k] = mid - half;
        bmax[k] = mid + half;
    }
    Root R;
    R.c[0] = (bmin[0] + bmax[0]) * 0.5f;
    R.c[1] = (bmin[1] + bmax[1]) * 0.5f;
    R.c[2] = (bmin[2] + bmax[2]) * 0.5f;
    R.c[3] = fmaxf(fmaxf(bmax[0] - bmin[0], bmax[1] - bmin[1]), bmax[2] - bmin[2]) * 0.5f * 1.001f;
    *out = R;
}

// Each block stores its min / max; the key pass, the next kernel, reduces
// the partials of its disks (no atomics, no fence: the kernel boundary
// orders the stores before the loads).
__device__ void bbox_post(float lx, float ly, float lz, float hx, float hy, float hz,
                          float *partial)
{
    block_minmax6(lx, ly, lz, hx, hy, hz, partial + 6 * blockIdx.x);
}

// Root cube of disk s from the block partials (BHTree.cu:440-446, :187-191).
// Called by every thread of the block; the result is valid in *out after
// the trailing barrier.
__device__ void root_of_partials(const SegTable &T, const float *__restrict__ partial, int s,
                                 Root *out)
{
    __shared__ float s_box[6];
    float a = FLT_MAX, b = FLT_MAX, c = FLT_MAX, d = -FLT_MAX, e = -FLT_MAX, f = -FLT_MAX;
    for (int q = T.blk[s] + threadIdx.x; q < T.blk[s + 1]; q += kBlock) {
        const float *p = partial + 6 * q;
        a = fminf(a, p[0]);
        b = fminf(b, p[1]);
        c = fminf(c, p[2]);
        d = fmaxf(d, p[3]);
        e = fmaxf(e, p[4]);
        f = fmaxf(f, p[5]);
    }
    block_minmax6(a, b, c, d, e, f, s_box);
    __syncthreads();
    if (threadIdx.x == 0) root_from_box(s_box, out);
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_bbox(SegTable T, const float *__restrict__ x,
                                                 const float *__restrict__ y,
                                                 const float *__restrict__ z,
                                                 float *partial)
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
    bbox_post(lx, ly, lz, hx, hy, hz, partial);
}

// ===========================================================================
// Keys: the octant digits of BHTree.cu:164-170 along the float descent of
// BHTree.cu:253-256, 21 levels, first level most significant.  The key pass
// also reduces the bounding-box partials to the root cubes and clears the
// look-back words and counters of the passes that follow.  For the radix
// sort it accumulates the digit histograms of all passes (each block covers
// kKeysPerBlock bodies, 8 per thread, so that few blocks post histograms);
// for the splitter sort it also performs the bucket scatter (below).
// ===========================================================================
constexpr int kKeysItems = 8;
constexpr int kKeysPerBlock = kBlock * kKeysItems;

__device__ __forceinline__ uint64_t morton_key(float px, float py, float pz, const Root &R0)
{
#pragma clang fp contract(off)
    float cx = R0.c[0], cy = R0.c[1], cz = R0.c[2], hh = R0.c[3];
    uint64_t key = 0;
#pragma unroll
    for (int l = 0; l < kLevels; ++l) {
        const bool ox = px > cx, oy = py > cy, oz = pz > cz;
        key = (key << 3) | (uint64_t)((ox ? 1 : 0) | (oy ? 2 : 0) | (oz ? 4 : 0));
        hh = hh * 0.5f;
        cx = cx + (ox ? hh : -hh);
        cy = cy + (oy ? hh : -hh);
        cz = cz + (oz ? hh : -hh);
    }
    return key;
}

// Root cubes of disks [s0, s1] into s_root (block 0: of every disk, also
// published to roots[]), and the resets owed by the key pass.
// nblocks: the blocks of the key pass (all call this).
__device__ void keys_prologue(const SegTable &T, const float *__restrict__ partial,
                              Root *__restrict__ roots, Root *s_root, int s0, int s1, int nblocks,
                              unsigned *__restrict__ status0, int n_status,
                              unsigned long long *__restrict__ scanw, int n_scanw,
                              int *__restrict__ tickets)
{
    if (blockIdx.x == 0) {
        s0 = 0;
        s1 = T.nseg - 1;
    }
    for (int sg = s0; sg <= s1; ++sg) root_of_partials(T, partial, sg, &s_root[sg]);
    if (blockIdx.x == 0 && threadIdx.x < T.nseg) roots[threadIdx.x] = s_root[threadIdx.x];
    const int gid = blockIdx.x * kBlock + threadIdx.x, gsz = nblocks * kBlock;
    for (int q = gid; q < n_status; q += gsz) status0[q] = 0;
    for (int q = gid; q < n_scanw; q += gsz) scanw[q] = kScanEmpty;
    if (gid <= kTicketScan2) tickets[gid] = 0;  // tickets of the radix passes and the scans
}

template <int BITS>
__global__ __launch_bounds__(kBlock) void k_keys(SegTable T, const float *__restrict__ x,
                                                 const float *__restrict__ y,
                                                 const float *__restrict__ z,
                                                 const float *__restrict__ partial,
                                                 Root *__restrict__ roots,
                                                 uint64_t *__restrict__ keys,
                                                 int *__restrict__ vals, int *__restrict__ ghist,
                                                 unsigned *__restrict__ status0, int n_status,
                                                 unsigned long long *__restrict__ scanw,
                                                 int n_scanw, int *__restrict__ tickets)
{
    constexpr int R = 1 << BITS, NP = (63 + BITS - 1) / BITS;
    __shared__ int h[NP * R];
    __shared__ Root s_root[kMaxSeg];
    if (ghist)
        for (int q = threadIdx.x; q < NP * R; q += kBlock) h[q] = 0;
    const int n = T.n_total;
    {
        const int b0 = blockIdx.x * kKeysPerBlock;
        const int b1 = min(n, b0 + kKeysPerBlock) - 1;
        keys_prologue(T, partial, roots, s_root, seg_of_body(T, b0), seg_of_body(T, b1),
                      gridDim.x, status0, n_status, scanw, n_scanw, tickets);
    }
    __syncthreads();
    float px[kKeysItems], py[kKeysItems], pz[kKeysItems];
#pragma unroll
    for (int r = 0; r < kKeysItems; ++r) {  // all loads first: one memory latency
        const int i = blockIdx.x * kKeysPerBlock + r * kBlock + threadIdx.x;
        px[r] = i < n ? x[i] : 0.f;
        py[r] = i < n ? y[i] : 0.f;
        pz[r] = i < n ? z[i] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < kKeysItems; ++r) {
        const int i = blockIdx.x * kKeysPerBlock + r * kBlock + threadIdx.x;
        if (i >= n) break;
        const uint64_t key = morton_key(px[r], py[r], pz[r], s_root[seg_of_body(T, i)]);
        keys[i] = key;
        if (ghist) {  // radix passes: index values and digit histograms
            vals[i] = i;
#pragma unroll
            for (int p = 0; p < NP; ++p)
                atomicAdd(&h[p * R + (int)((key >> (BITS * p)) & (uint64_t)(R - 1))], 1);
        }
    }
    if (!ghist) return;
    __syncthreads();
    for (int q = threadIdx.x; q < NP * R; q += kBlock)
        if (h[q]) atomicAdd(&ghist[q], h[q]);
}

// ===========================================================================
// One stable radix pass (BITS-bit digit) with decoupled look-back.  A tile
// of kSweepTile keys takes a ticket in launch order; each of its four waves
// ranks its own contiguous quarter in rounds of 64 keys (wave match with
// BITS ballots, per-wave running digit counters in LDS, no block barrier);
// the waves' counts are then combined per digit, posted, and the counts of
// earlier tiles are read back through one status word per (tile, digit):
// bit 31 inclusive prefix, bit 30 tile aggregate, low 30 bits the count.
// mode 0: digit = key bits [BITS pass, BITS pass + BITS);
// mode 1: digit = disk of the body (the final pass of a batched sort).
// ===========================================================================
constexpr unsigned kAgg = 1u << 30, kInc = 1u << 31, kValMask = kAgg - 1;

// Exclusive scan over the 256 threads of a block; returns the exclusive
// prefix of v, *total receives the sum.
__device__ inline int block_excl_scan256(int v, int *total)
{
    __shared__ int s_w[kBlock / kWave];
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kBlock / kWave; ++k) {
        const int c = s_w[k];
        if (k < w) before += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return before + inc - v;
}


template <int BITS>
__global__ __launch_bounds__(kBlock) void k_sweep(
    const uint64_t *__restrict__ kin, const int *__restrict__ vin, uint64_t *__restrict__ kout,
    int *__restrict__ vout, int n, int pass, int mode, SegTable T, const int *__restrict__ ghist,
    unsigned *status, unsigned *status_next, int *tickets, int gather,
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
    const float *__restrict__ m, float4 *__restrict__ bodies)
{
    constexpr int R = 1 << BITS;
    constexpr int DPT = R / kBlock;              // digits owned by each thread (>= 1)
    constexpr int ROUNDS = kSweepTile / kBlock;  // rounds of 64 keys per wave
    // look-back window (tiles per round of independent loads): all tiles post
    // their aggregates at about the same time, so a wide window resolves the
    // start-up chain in a few memory round trips
    constexpr int LB = DPT == 1 ? 32 : (DPT == 2 ? 16 : 4);
    static_assert(DPT >= 1 && R % kBlock == 0, "radix must be a multiple of the block");
    __shared__ int s_cnt[kBlock / kWave][R];
    __shared__ int s_goff[R];
    __shared__ int s_tile;
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    if (t == 0) s_tile = atomicAdd(&tickets[kTicketPass + pass], 1);
    // global exclusive digit offsets: thread t owns digits [t DPT, (t+1) DPT)
    int g[DPT];
    int gs = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int d = t * DPT + q;
        g[q] = mode == 0 ? ghist[pass * R + d] : (d < T.nseg ? T.off[d + 1] - T.off[d] : 0);
        gs += g[q];
    }
    int total;
    int run = block_excl_scan256(gs, &total);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        s_goff[t * DPT + q] = run;
        run += g[q];
    }
#pragma unroll
    for (int q = 0; q < kBlock / kWave; ++q)
#pragma unroll
        for (int u = 0; u < DPT; ++u) s_cnt[q][t * DPT + u] = 0;
    __syncthreads();
    const int tile = s_tile;
    if (status_next)
#pragma unroll
        for (int q = 0; q < DPT; ++q) status_next[tile * R + t * DPT + q] = 0;
    const int shift = BITS * pass;
    const int base = tile * kSweepTile + w * (kSweepTile / (kBlock / kWave));
    uint64_t kk[ROUNDS];
    int vv[ROUNDS], dd[ROUNDS], lo[ROUNDS];
    int *cnt = s_cnt[w];
    // issue every load of the wave's chunk before ranking: one memory latency
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int i = base + r * kWave + lane;
        kk[r] = i < n ? kin[i] : 0ull;
        vv[r] = i < n ? vin[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int i = base + r * kWave + lane;
        const bool valid = i < n;
        const uint64_t key = kk[r];
        const int val = vv[r];
        int d = 0;
        if (valid) d = mode == 0 ? (int)((key >> shift) & (uint64_t)(R - 1)) : seg_of_body(T, val);
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int bit = 0; bit < BITS; ++bit) {
            const bool set = (d >> bit) & 1;
            const uint64_t bb = ballot64(set);
            peers &= set ? bb : ~bb;
        }
        const int rank = __popcll(peers & lanemask_lt());
        const int before = cnt[d];  // wave-private counter: program order suffices
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) cnt[d] = before + __popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        dd[r] = valid ? d : -1;
        lo[r] = before + rank;
    }
    __syncthreads();
    // per digit: exclusive offsets of the waves inside the tile, tile count
    unsigned tc[DPT];
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int d = t * DPT + q;
        const int c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        s_cnt[0][d] = 0;
        s_cnt[1][d] = c0;
        s_cnt[2][d] = c0 + c1;
        s_cnt[3][d] = c0 + c1 + c2;
        tc[q] = (unsigned)(c0 + c1 + c2 + c3);
    }
    // post this tile's counts, then look back over earlier tiles, LB tiles
    // and all owned digits per round of independent loads
    unsigned pre[DPT];
    bool done[DPT];
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        pre[q] = 0;
        done[q] = tile == 0;
        __hip_atomic_store(status + tile * R + t * DPT + q, (tile == 0 ? kInc : kAgg) | tc[q],
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (int j = tile - 1; j >= 0; j -= LB) {
        bool all = true;
#pragma unroll
        for (int q = 0; q < DPT; ++q) all = all && done[q];
        if (all) break;
        unsigned v[DPT][LB];
#pragma unroll
        for (int q = 0; q < DPT; ++q)
#pragma unroll
            for (int u = 0; u < LB; ++u)
                v[q][u] = (!done[q] && j - u >= 0)
                              ? __hip_atomic_load(status + (j - u) * R + t * DPT + q,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : kInc;
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                if (done[q]) break;
                unsigned val = v[q][u];
                while (val == 0) {  // not posted yet: wait for that tile
                    __builtin_amdgcn_s_sleep(1);
                    val = __hip_atomic_load(status + (j - u) * R + t * DPT + q, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
                }
                pre[q] += val & kValMask;
                if (val & kInc) done[q] = true;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        if (tile > 0)
            __hip_atomic_store(status + tile * R + t * DPT + q, kInc | (pre[q] + tc[q]),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_goff[t * DPT + q] += (int)pre[q];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int d = dd[r];
        if (d < 0) continue;
        const int pos = s_goff[d] + s_cnt[w][d] + lo[r];
        kout[pos] = kk[r];
        vout[pos] = vv[r];
        if (gather) {
            const int v = vv[r];
            bodies[pos] = make_float4(x[v], y[v], z[v], m[v]);
        }
    }
}

// ===========================================================================
// Splitter-bucket sort.  The sorted order is the ascending order of the
// pairs (key, body index): the order the stable radix passes above produce
// from index-ordered input, so either sort feeds the same tree.
//   splitters      bucket b of a disk holds the pairs from its lower
//                  splitter up to the next one.  Every bucket sort writes,
//                  as its splitters for the next build, the pairs at the
//                  quantile positions floor(b n / buckets) of the order it
//                  produced (bodies move little between steps, so buckets
//                  stay near kBucketMean members).  A build without them
//                  (the first of a tree, and every build outside the step
//                  loop) samples kOversample keys per bucket (k_split).
//   scatter        blocks of kSplitTile bodies of one disk find each body's
//                  bucket by binary search over the splitters in LDS, group
//                  the block's bodies by bucket (LDS counters and a block
//                  scan) and store them at the block's own body range, with
//                  one packed {offset << 16 | count} word per (bucket,
//                  block).  No global atomics.  In the step loop the scatter
//                  runs inside the key pass (k_keys_scatter).
//   k_bucket_sort  one block per bucket gathers the bucket's members through
//                  its column of words, sorts them with an all-ascending
//                  bitonic network in LDS (a bucket beyond kBucketCap
//                  members, possible only when the splitters misrepresent
//                  the keys, is sorted in place in global memory) and writes
//                  keys, indices and the gathered {x, y, z, m} at the
//                  bucket's offset: the sum of the in-block offsets of its
//                  words.
// Every (key, index) pair is distinct, so the result depends neither on the
// splitters nor on the order in which LDS atomics rank a bucket's members.
// ===========================================================================
__device__ __forceinline__ bool kv_less(uint64_t ka, int ia, uint64_t kb, int ib)
{
    return ka < kb || (ka == kb && ia < ib);
}

constexpr uint64_t kPadKey = ~0ull;  // above every 63-bit key
constexpr int kPadIdx = 0x7fffffff;

// Exclusive scan over the NT threads of a block; *total receives the sum.
template <int NT>
__device__ inline int block_excl_scan(int v, int *total)
{
    __shared__ int s_w[NT / kWave];
    const int lane = __lane_id(), w = threadIdx.x / kWave;
    int inc = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) {
        const int c = s_w[k];
        if (k < w) before += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return before + inc - v;
}

// One comparator stage of the all-ascending bitonic network on one element
// per lane: partners are lane and lane ^ mask, hib is the top bit of mask,
// and the lower lane of every pair keeps the smaller element.  Elements are
// distinct except padding, and equal elements may be exchanged freely, so
// the upper lane takes the partner exactly when the lower one does not.
// The value of lane (lane ^ mask).  Partners inside a row of 16 lanes come
// over DPP at ALU speed (quad permutes, the row and half-row mirrors, row
// shifts by 4 and 8 selected by the lane's bit); the others over the lane
// shuffle, which goes through the LDS crossbar.  After unrolling, mask is a
// constant and the switch folds away.
__device__ __forceinline__ int xor_lane(int v, int mask, int lane)
{
    switch (mask) {
    case 1: return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    case 2: return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    case 3: return __builtin_amdgcn_update_dpp(v, v, 0x1B, 0xf, 0xf, false);   // quad_perm [3,2,1,0]
    case 7: return __builtin_amdgcn_update_dpp(v, v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    case 15: return __builtin_amdgcn_update_dpp(v, v, 0x140, 0xf, 0xf, false); // row_mirror
    case 4: {
        const int up = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xf, 0xf, false);  // row_shl:4
        const int dn = __builtin_amdgcn_update_dpp(v, v, 0x114, 0xf, 0xf, false);  // row_shr:4
        return (lane & 4) ? dn : up;
    }
    case 8: {
        const int up = __builtin_amdgcn_update_dpp(v, v, 0x108, 0xf, 0xf, false);  // row_shl:8
        const int dn = __builtin_amdgcn_update_dpp(v, v, 0x118, 0xf, 0xf, false);  // row_shr:8
        return (lane & 8) ? dn : up;
    }
    default: return __shfl_xor(v, mask, kWave);
    }
}

__device__ __forceinline__ void wave_cmpx(uint64_t &k, int &id, int lane, int mask, int hib)
{
    const unsigned plo = (unsigned)xor_lane((int)(unsigned)k, mask, lane);
    const unsigned phi = (unsigned)xor_lane((int)(unsigned)(k >> 32), mask, lane);
    const int pid = xor_lane(id, mask, lane);
    const uint64_t pk = ((uint64_t)phi << 32) | plo;
    if (kv_less(pk, pid, k, id) == ((lane & hib) == 0)) {
        k = pk;
        id = pid;
    }
}

template <class KP, class IP>
__device__ __forceinline__ void cmpx_mem(KP sk, IP si, int i, int l)
{
    const uint64_t a = sk[i], b = sk[l];
    const int ia = si[i], ib = si[l];
    if (kv_less(b, ib, a, ia)) {
        sk[i] = b;
        si[i] = ib;
        sk[l] = a;
        si[l] = ia;
    }
}

// Sorts sk/si[0, c) ascending by (key, index) with all NT threads of the
// block.  The network is the bitonic sorter in its all-ascending form (each
// merge starts with the mirrored comparator i <-> i ^ (k - 1)), so the
// virtual padding of [c, 2^ceil(log2 c)) with +inf never moves and every
// comparator that touches it is skipped.  Stages within 64 elements run in
// registers (lane shuffles).  sk/si are the block's own LDS or global
// memory (separate instantiations, so that LDS stays on ds_* accesses).
template <int NT, class KP, class IP>
__device__ __forceinline__ void block_sort(KP sk, IP si, int c)
{
    constexpr int NW = NT / kWave;
    const int lane = __lane_id(), w = threadIdx.x / kWave;
    const int nch = (c + kWave - 1) / kWave;
    for (int ch = w; ch < nch; ch += NW) {
        const int i = ch * kWave + lane;
        uint64_t k = i < c ? sk[i] : kPadKey;
        int id = i < c ? si[i] : kPadIdx;
#pragma unroll
        for (int kk = 2; kk <= kWave; kk <<= 1) {
            wave_cmpx(k, id, lane, kk - 1, kk >> 1);
#pragma unroll
            for (int j = kk >> 2; j > 0; j >>= 1) wave_cmpx(k, id, lane, j, j);
        }
        if (i < c) {
            sk[i] = k;
            si[i] = id;
        }
    }
    __syncthreads();
    int P = kWave;
    while (P < c) P <<= 1;
    for (int kk = 2 * kWave; kk <= P; kk <<= 1) {
        const int half = kk >> 1;
        for (int q = threadIdx.x; q < P / 2; q += NT) {
            const int i = q + (q & ~(half - 1));  // lower half of its kk-block
            const int l = i ^ (kk - 1);
            if (l < c) cmpx_mem(sk, si, i, l);
        }
        __syncthreads();
        for (int j = kk >> 2; j >= kWave; j >>= 1) {
            for (int q = threadIdx.x; q < P / 2; q += NT) {
                const int i = q + (q & ~(j - 1));
                const int l = i + j;
                if (l < c) cmpx_mem(sk, si, i, l);
            }
            __syncthreads();
        }
        for (int ch = w; ch < nch; ch += NW) {
            const int i = ch * kWave + lane;
            uint64_t k = i < c ? sk[i] : kPadKey;
            int id = i < c ? si[i] : kPadIdx;
#pragma unroll
            for (int j = kWave / 2; j > 0; j >>= 1) wave_cmpx(k, id, lane, j, j);
            if (i < c) {
                sk[i] = k;
                si[i] = id;
            }
        }
        __syncthreads();
    }
}

constexpr int kSplitThreads = 1024;

__global__ __launch_bounds__(kSplitThreads) void k_split(SegTable T,
                                                         const uint64_t *__restrict__ keys,
                                                         uint64_t *__restrict__ spk,
                                                         int *__restrict__ spi)
{
    __shared__ uint64_t sk[kMaxSample];
    __shared__ int si[kMaxSample];
    const int s = blockIdx.x;
    const int n = T.off[s + 1] - T.off[s];
    const int nb = T.bkoff[s + 1] - T.bkoff[s];
    if (nb <= 1) return;
    const int S = min(n, kOversample * nb);
    const double stride = (double)n / (double)S;  // >= 1: distinct indices
    for (int j = threadIdx.x; j < S; j += kSplitThreads) {
        const int idx = T.off[s] + (int)(((double)j + 0.5) * stride);
        sk[j] = keys[idx];
        si[j] = idx;
    }
    __syncthreads();
    block_sort<kSplitThreads>(sk, si, S);
    for (int b = 1 + threadIdx.x; b < nb; b += kSplitThreads) {
        const int q = (int)((long long)b * S / nb);
        spk[T.bkoff[s] + b] = sk[q];
        spi[T.bkoff[s] + b] = si[q];
    }
}

// Scatter of one block of kSplitTile bodies of disk s (block kb of the
// disk): kk[r] is the key of body base + r kBlock + threadIdx.x (valid below
// end).  All threads of the block call it.
__device__ __forceinline__ void scatter_tile(const SegTable &T, int s, int kb, int base, int end,
                                             const uint64_t (&kk)[kSplitItems],
                                             const float4 (&pay)[kSplitItems],
                                             const uint64_t *__restrict__ spk,
                                             const int *__restrict__ spi,
                                             uint64_t *__restrict__ skey,
                                             int *__restrict__ sidx, float4 *__restrict__ spay,
                                             int *__restrict__ bm)
{
    __shared__ uint64_t s_k[kMaxBuckets];
    __shared__ int s_i[kMaxBuckets];
    __shared__ int s_c[kMaxBuckets];
    __shared__ int s_o[kMaxBuckets];
    constexpr int BPT = kMaxBuckets / kBlock;  // buckets per thread in the count scan
    const int t = threadIdx.x;
    const int nb = T.bkoff[s + 1] - T.bkoff[s];
    const int nblk = T.sblk[s + 1] - T.sblk[s];
    for (int q = t; q < nb; q += kBlock) {
        s_c[q] = 0;
        if (q > 0) {
            s_k[q] = spk[T.bkoff[s] + q];
            s_i[q] = spi[T.bkoff[s] + q];
        }
    }
    __syncthreads();
    int top = 1;
    while (2 * top < nb) top <<= 1;
    int bk[kSplitItems], lr[kSplitItems];
#pragma unroll
    for (int r = 0; r < kSplitItems; ++r) {
        const int i = base + r * kBlock + t;
        bk[r] = -1;
        lr[r] = 0;
        if (i >= end) continue;
        // largest bucket whose lower splitter is <= (key, i); bucket 0 has none
        int lo = 0;
        for (int step = top; step > 0; step >>= 1) {
            const int cand = lo + step;
            if (cand < nb && !kv_less(kk[r], i, s_k[cand], s_i[cand])) lo = cand;
        }
        bk[r] = lo;
        lr[r] = atomicAdd(&s_c[lo], 1);
    }
    __syncthreads();
    // exclusive offsets of the buckets inside the block
    int c[BPT], sum = 0;
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int q = t * BPT + u;
        c[u] = q < nb ? s_c[q] : 0;
        sum += c[u];
    }
    int total;
    int run = block_excl_scan<kBlock>(sum, &total);
    int *col = bm + T.moff[s] + kb;
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int q = t * BPT + u;
        if (q < nb) {
            s_o[q] = run;
            col[q * nblk] = (run << 16) | c[u];
        }
        run += c[u];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSplitItems; ++r) {
        if (bk[r] < 0) continue;
        const int pos = base + s_o[bk[r]] + lr[r];
        skey[pos] = kk[r];
        sidx[pos] = base + r * kBlock + t;
        spay[pos] = pay[r];
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(
    SegTable T, const uint64_t *__restrict__ keys, const float *__restrict__ x,
    const float *__restrict__ y, const float *__restrict__ z, const float *__restrict__ m,
    const uint64_t *__restrict__ spk, const int *__restrict__ spi, uint64_t *__restrict__ skey,
    int *__restrict__ sidx, float4 *__restrict__ spay, int *__restrict__ bm)
{
    const int s = seg_of_sblock(T, blockIdx.x);
    const int kb = blockIdx.x - T.sblk[s];
    const int base = T.off[s] + kb * kSplitTile;
    const int end = min(T.off[s + 1], base + kSplitTile);
    uint64_t kk[kSplitItems];
    float4 pay[kSplitItems];
#pragma unroll
    for (int r = 0; r < kSplitItems; ++r) {
        const int i = base + r * kBlock + threadIdx.x;
        kk[r] = i < end ? keys[i] : 0ull;
        pay[r] = i < end ? make_float4(x[i], y[i], z[i], m[i]) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    scatter_tile(T, s, kb, base, end, kk, pay, spk, spi, skey, sidx, spay, bm);
}

// Key pass and scatter in one kernel (splitters of the previous build); the
// blocks past the scatter blocks place the next walk's wave order.
__global__ __launch_bounds__(kBlock) void k_keys_scatter(
    SegTable T, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ z, const float *__restrict__ m, const float *__restrict__ partial,
    Root *__restrict__ roots, unsigned long long *__restrict__ scanw, int n_scanw,
    int *__restrict__ tickets, const uint64_t *__restrict__ spk,
    const int *__restrict__ spi, uint64_t *__restrict__ skey, int *__restrict__ sidx,
    float4 *__restrict__ spay, int *__restrict__ bm, LptDev L)
{
    __shared__ Root s_root[kMaxSeg];
    if (blockIdx.x >= T.sblk[T.nseg]) {  // uniform per block
        lpt_scatter<kBlock>(L, blockIdx.x - T.sblk[T.nseg]);
        return;
    }
    const int s = seg_of_sblock(T, blockIdx.x);
    const int kb = blockIdx.x - T.sblk[s];
    const int base = T.off[s] + kb * kSplitTile;
    const int end = min(T.off[s + 1], base + kSplitTile);
    float px[kSplitItems], py[kSplitItems], pz[kSplitItems], pm[kSplitItems];
#pragma unroll
    for (int r = 0; r < kSplitItems; ++r) {
        const int i = base + r * kBlock + threadIdx.x;
        px[r] = i < end ? x[i] : 0.f;
        py[r] = i < end ? y[i] : 0.f;
        pz[r] = i < end ? z[i] : 0.f;
        pm[r] = i < end ? m[i] : 0.f;
    }
    keys_prologue(T, partial, roots, s_root, s, s, T.sblk[T.nseg], nullptr, 0, scanw, n_scanw,
                  tickets);
    __syncthreads();
    uint64_t kk[kSplitItems];
    float4 pay[kSplitItems];
#pragma unroll
    for (int r = 0; r < kSplitItems; ++r) {
        kk[r] = morton_key(px[r], py[r], pz[r], s_root[s]);
        pay[r] = make_float4(px[r], py[r], pz[r], pm[r]);
    }
    scatter_tile(T, s, kb, base, end, kk, pay, spk, spi, skey, sidx, spay, bm);
}

constexpr int kSortThreads = 512;

// Staging position of bucket member j: contribution k with
// s_pre[k] <= j < s_pre[k + 1] (top: the largest power of two <= nblk).
__device__ __forceinline__ int member_src(const int *s_pre, const int *s_src, int nblk, int top,
                                          int j)
{
    int k = 0;
    for (int step = top; step > 0; step >>= 1)
        if (k + step < nblk && s_pre[k + step] <= j) k += step;
    return s_src[k] + (j - s_pre[k]);
}

// Splitter of the next build at disk position p, if p is a quantile
// position floor(b n / nb) (p nb and b n stay below 2^31: n <= kSplitMaxN,
// nb <= kMaxBuckets).
__device__ __forceinline__ void emit_splitter(const SegTable &T, int s, int p, uint64_t key, int v,
                                              uint64_t *__restrict__ spk, int *__restrict__ spi)
{
    const int n = T.off[s + 1] - T.off[s];
    const int nb = T.bkoff[s + 1] - T.bkoff[s];
    const int q = (p * nb + n - 1) / n;
    if (q >= 1 && q < nb && q * n / nb == p) {
        spk[T.bkoff[s] + q] = key;
        spi[T.bkoff[s] + q] = v;
    }
}

// The bucket sort's sort value in LDS: (body index - disk offset) << 11 |
// member slot.  Ordering by it is ordering by body index (n_s <= 2^20), and
// the slot finds the member's staged {x, y, z, m} after sorting (c <= 2048).
constexpr int kSlotBits = 11;
static_assert(kBucketCap <= (1 << kSlotBits) && kSplitMaxN <= (1 << (31 - kSlotBits)),
              "index and slot share 31 bits");

__global__ __launch_bounds__(kSortThreads) void k_bucket_sort(
    SegTable T, const uint64_t *__restrict__ skey, const int *__restrict__ sidx,
    const float4 *__restrict__ spay, const int *__restrict__ bm, uint64_t *__restrict__ kout,
    int *__restrict__ vout, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ z, const float *__restrict__ m, float4 *__restrict__ bodies,
    uint64_t *__restrict__ spk, int *__restrict__ spi, int nblk_max)
{
    __shared__ uint64_t sk[kBucketCap];
    __shared__ int si[kBucketCap];
    // per contribution (scatter block of the disk): the first bucket member
    // and the staging position, sized at launch by the most scatter blocks
    // of any disk of the tree (nblk_max + 1 and nblk_max ints)
    extern __shared__ int s_dyn[];
    int *s_pre = s_dyn, *s_src = s_dyn + nblk_max + 1;
    constexpr int CPT = kMaxSplitBlocks / kSortThreads;
    const int t = threadIdx.x;
    const int g = blockIdx.x;
    const int s = seg_of_bucket(T, g);
    const int b = g - T.bkoff[s];
    const int nblk = T.sblk[s + 1] - T.sblk[s];
    const int *row = bm + T.moff[s] + b * nblk;
    int cw[CPT], csum = 0, osum = 0;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int k = t * CPT + u;
        cw[u] = k < nblk ? row[k] : 0;
        csum += cw[u] & 0xffff;
        osum += cw[u] >> 16;
    }
    int c, earlier;
    int run = block_excl_scan<kSortThreads>(csum, &c);
    // members of disk s in earlier buckets: the sum over blocks of the offsets
    (void)block_excl_scan<kSortThreads>(osum, &earlier);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int k = t * CPT + u;
        if (k < nblk) {
            s_pre[k] = run;
            s_src[k] = T.off[s] + k * kSplitTile + (cw[u] >> 16);
        }
        run += cw[u] & 0xffff;
    }
    __syncthreads();
    int top = 1;
    while (2 * top <= nblk) top <<= 1;
    const int off = T.off[s], out = off + earlier;
    if (c <= kBucketCap) {
        for (int j = t; j < c; j += kSortThreads) {
            const int q = member_src(s_pre, s_src, nblk, top, j);
            sk[j] = skey[q];
            si[j] = ((sidx[q] - off) << kSlotBits) | j;
        }
        __syncthreads();
        block_sort<kSortThreads>(sk, si, c);
        for (int j = t; j < c; j += kSortThreads) {
            const int w = si[j];
            const int v = off + (w >> kSlotBits);
            const uint64_t key = sk[j];
            kout[out + j] = key;
            vout[out + j] = v;
            bodies[out + j] = spay[member_src(s_pre, s_src, nblk, top, w & ((1 << kSlotBits) - 1))];
            emit_splitter(T, s, earlier + j, key, v, spk, spi);
        }
    } else {  // in place at the bucket's final range, by body index
        uint64_t *ok = kout + out;
        int *ov = vout + out;
        for (int j = t; j < c; j += kSortThreads) {
            const int q = member_src(s_pre, s_src, nblk, top, j);
            ok[j] = skey[q];
            ov[j] = sidx[q];
        }
        __syncthreads();
        block_sort<kSortThreads>(ok, ov, c);
        for (int j = t; j < c; j += kSortThreads) {
            const int v = ov[j];
            bodies[out + j] = make_float4(x[v], y[v], z[v], m[v]);
            emit_splitter(T, s, earlier + j, ok[j], v, spk, spi);
        }
    }
}

// ===========================================================================
// Exclusive prefix P[i] = sum_{s<i} e(s) of per-body moment vectors of the
// sorted bodies in double precision (single pass; tile prefixes summed in a
// fixed order, see below):
//   KIND 0: e = {m, m x, m y, m z}            (mass and COM of every node)
//   KIND 1: e = {m xx, m yy, m zz, m xy, m xz, m yz}   (quadrupoles)
// A node's moments are differences over its range of sorted positions.
// ===========================================================================
template <int N>
__device__ inline DVec<N> dv_add(const DVec<N> &a, const DVec<N> &b)
{
    DVec<N> r;
#pragma unroll
    for (int q = 0; q < N; ++q) r.v[q] = a.v[q] + b.v[q];
    return r;
}

// Look-back words: one relaxed 64-bit atomic per double (no fence needed:
// a reader only uses a value whose N words all differ from kScanEmpty).
template <int N>
__device__ inline void dv_publish(DVec<N> *p, const DVec<N> &a)
{
#pragma unroll
    for (int q = 0; q < N; ++q)
        __hip_atomic_store((unsigned long long *)&p->v[q],
                           (unsigned long long)__double_as_longlong(a.v[q]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
template <int N>
__device__ inline bool dv_fetch(const DVec<N> *p, DVec<N> &a)
{
    bool ok = true;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const unsigned long long u = __hip_atomic_load((unsigned long long *)&p->v[q],
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = ok && u != kScanEmpty;
        a.v[q] = __longlong_as_double((long long)u);
    }
    return ok;
}

template <int N>
__device__ inline DVec<N> dv_zero()
{
    DVec<N> r;
#pragma unroll
    for (int q = 0; q < N; ++q) r.v[q] = 0.0;
    return r;
}

template <int N>
__device__ inline DVec<N> dv_shfl_up(const DVec<N> &a, int o)
{
    DVec<N> r;
#pragma unroll
    for (int q = 0; q < N; ++q) r.v[q] = __shfl_up(a.v[q], o, 64);
    return r;
}

// Disk s owns prefix entries [off[s] + s, off[s+1] + s]: its own zero and
// its n_s running sums, so every disk's sums start from zero exactly as in a
// plain run of that disk alone.
__host__ __device__ inline int pbase(const SegTable &T, int s) { return T.off[s] + s; }

// KIND 0: {m, m x, m y, m z}; KIND 1: second moments; KIND 2: KIND 0 with
// the positions of the FP64 state s64 ([x | y | z] in body order, read
// through the sorted order perm) instead of the float bodies.
template <int N, int KIND>
__device__ inline void prefix_tile(const SegTable &T, const float4 *__restrict__ bodies,
                                   DVec<N> *P, DVec<N> *pagg, DVec<N> *pinc, int *tickets,
                                   int slot, int *ghist, const double *__restrict__ s64,
                                   const int *__restrict__ perm)
{
    __shared__ int s_tile;
    __shared__ DVec<N> s_w[kBlock / kWave];
    __shared__ DVec<N> s_pre;
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    if (t == 0) s_tile = atomicAdd(&tickets[slot], 1);
    __syncthreads();
    const int tile = s_tile;
    const int s = seg_of_stile(T, tile);
    const int first_tile = T.stile[s];
    const int ns = T.off[s + 1] - T.off[s];
    const int lbase = (tile - first_tile) * kScanTile + t * kScanItems;  // disk-local
    DVec<N> v[kScanItems];
    DVec<N> sum = dv_zero<N>();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = lbase + k;
        DVec<N> e = dv_zero<N>();
        if (i < ns) {
            const float4 b = bodies[T.off[s] + i];
            const double mm = b.w;
            double x = b.x, y = b.y, z = b.z;
            if constexpr (KIND == 2) {
                const int n = T.n_total, j = perm[T.off[s] + i];
                x = s64[j];
                y = s64[n + j];
                z = s64[2 * n + j];
            }
            if constexpr (KIND != 1) {
                e.v[0] = mm;
                e.v[1] = mm * x;
                e.v[2] = mm * y;
                e.v[3] = mm * z;
            } else {
                e.v[0] = mm * x * x;
                e.v[1] = mm * y * y;
                e.v[2] = mm * z * z;
                e.v[3] = mm * x * y;
                e.v[4] = mm * x * z;
                e.v[5] = mm * y * z;
            }
        }
        v[k] = e;
        sum = dv_add(sum, e);
    }
    DVec<N> inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const DVec<N> y = dv_shfl_up(inc, o);
        if (lane >= o) inc = dv_add(inc, y);
    }
    DVec<N> excl = dv_shfl_up(inc, 1);
    if (lane == 0) excl = dv_zero<N>();
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    DVec<N> before = dv_zero<N>(), agg = dv_zero<N>();
#pragma unroll
    for (int k = 0; k < kBlock / kWave; ++k) {
        if (k < w) before = dv_add(before, s_w[k]);
        agg = dv_add(agg, s_w[k]);
    }
    if (w == 0) {
        // Prefix of the tile aggregates with a summation order that never
        // depends on timing (so a rebuild from the same bodies gives the same
        // COMs bit for bit).  The tiles of a disk form blocks of 64; the last
        // tile of a block sums the block's aggregates in a fixed order and
        // posts the sum at the block's first tile; a tile adds the sums of
        // the earlier blocks and the aggregates of the earlier tiles of its
        // own block.  Words are 64-bit relaxed atomics, complete when none is
        // kScanEmpty; every wait is on a tile with an earlier ticket.
        const int tl = tile - first_tile, g0 = first_tile + (tl & ~63);
        const int last_tile = T.stile[s + 1] - 1;
        auto wait_fetch = [&](const DVec<N> *p) {
            DVec<N> v;
            while (!dv_fetch(p, v)) __builtin_amdgcn_s_sleep(1);
            return v;
        };
        auto wave_sum = [&](DVec<N> c) {
#pragma unroll
            for (int q = 0; q < N; ++q)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) c.v[q] += __shfl_xor(c.v[q], o, 64);
            return c;
        };
        if (lane == 0) dv_publish(pagg + tile, agg);
        if ((tl & 63) == 63 || tile == last_tile) {  // close this block
            const int j = g0 + lane;
            DVec<N> c = dv_zero<N>();
            if (j < tile) c = wait_fetch(pagg + j);
            else if (j == tile) c = agg;
            c = wave_sum(c);
            if (lane == 0) dv_publish(pinc + g0, c);
        }
        DVec<N> pre = dv_zero<N>();
        for (int gb = first_tile; gb < g0; gb += 64 * 64) {  // earlier blocks, 64 at a time
            const int jb = gb + 64 * lane;
            DVec<N> c = dv_zero<N>();
            if (jb < g0) c = wait_fetch(pinc + jb);
            pre = dv_add(pre, wave_sum(c));
        }
        DVec<N> c = dv_zero<N>();  // earlier tiles of this block
        if (g0 + lane < tile) c = wait_fetch(pagg + g0 + lane);
        pre = dv_add(pre, wave_sum(c));
        if (lane == 0) s_pre = pre;
        if (tile == first_tile && lane == 0) P[pbase(T, s)] = dv_zero<N>();
    }
    if (KIND != 1 && tile == 0)  // histograms are consumed: clear them for the next build
        for (int q = t; q < kHistInts; q += kBlock) ghist[q] = 0;
    __syncthreads();
    DVec<N> run = dv_add(s_pre, dv_add(before, excl));
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        run = dv_add(run, v[k]);
        if (lbase + k < ns) P[pbase(T, s) + lbase + k + 1] = run;
    }
}

// ===========================================================================
// Binary radix tree over the sorted keys of each disk (longest common
// prefix search, ties broken by position).  Internal node j of disk s covers
// sorted positions [lo, hi]; children are encoded >= 0 (internal, global
// index) or < 0 (leaf, ~global sorted position).  The same thread derives
// the node's octree level floor(prefix bits / 3), whether it starts a new
// octree cell (its parent's prefix is max(lcp(lo-1, lo), lcp(hi, hi+1))),
// and the cell's centre by the float descent of the keys.  It needs only the
// sorted keys, so it runs in the launch of the moment prefix (below); the
// record pass takes each node's mass and COM from the prefix.
// ===========================================================================

__device__ inline int lcp(const uint64_t *__restrict__ k, int n, int a, int b)
{
    if (a < 0 || b < 0 || a >= n || b >= n) return -1;
    const uint64_t ka = k[a], kb = k[b];
    if (ka == kb) return 64 + __clz((unsigned)(a ^ b));
    return __clzll(ka ^ kb);
}

// lcp(k[i], k[b]) for k[i] = ki from the key kb read at b's clamped index;
// -1 for b outside [0, n).
__device__ inline int lcp_kb(int n, int i, uint64_t ki, int b, uint64_t kb)
{
    const int v = ki == kb ? 64 + __clz((unsigned)(i ^ b)) : __clzll(ki ^ kb);
    return (b < 0 || b >= n) ? -1 : v;
}
__device__ inline uint64_t key_at(const uint64_t *__restrict__ k, int n, int b)
{
    return k[min(max(b, 0), n - 1)];
}
__device__ inline int lcp_i(const uint64_t *__restrict__ k, int n, int i, uint64_t ki, int b)
{
    return lcp_kb(n, i, ki, b, key_at(k, n, b));
}

// Last m in [0, span) with lcp(i, i + (base + m) d) > thr, a predicate that
// holds at m = 0 and, once false, stays false.  Three bits of m per round:
// the 7 probe keys of a round are read together (the scheduling barrier
// keeps every load ahead of the first compare), so a search over 2^e
// offsets costs ceil(e / 3) round trips instead of e.
__device__ inline int last_true(const uint64_t *__restrict__ k, int n, int i, uint64_t ki, int d,
                                int base, int thr, int span)
{
    int o = 0;
    if (span <= 1) return 0;
    for (int b = 31 - __clz(span - 1); b >= 0; b -= 3) {
        const int low = max(b - 2, 0), ncand = (2 << (b - low)) - 1, step = 1 << low;
        uint64_t kb[7];
#pragma unroll
        for (int c = 1; c <= 7; ++c) kb[c - 1] = key_at(k, n, i + (base + o + c * step) * d);
        __builtin_amdgcn_sched_barrier(0);
        int cnt = 0;  // the probes that hold come first
#pragma unroll
        for (int c = 1; c <= 7; ++c) {
            const int m = o + c * step, pos = i + (base + m) * d;
            cnt += ((c <= ncand) & (m < span) & (lcp_kb(n, i, ki, pos, kb[c - 1]) > thr)) ? 1 : 0;
        }
        o += cnt * step;
    }
    return o;
}

// Octree level of a binary node: floor(prefix bits / 3); identical keys
// (prefix >= 64) form level-21 cells.  Key bit 63 is always zero.
__device__ inline int oct_level(int plen) { return plen >= 64 ? kLevels : (plen - 1) / 3; }

__device__ inline void karras_node(const SegTable &T, const uint64_t *__restrict__ keys,
                                   const Root *__restrict__ roots, int *left, int *right,
                                   int *plen, int *octf, float4 *bin_cell, int2 *brange, int j)
{
#pragma clang fp contract(off)
    if (j >= T.ioff[T.nseg]) return;
    const int s = seg_of_internal(T, j);
    const int n = T.off[s + 1] - T.off[s];
    const uint64_t *k = keys + T.off[s];
    const int i = j - T.ioff[s];
    const uint64_t ki = k[i];
    const int lp = lcp_i(k, n, i, ki, i + 1), lm = lcp_i(k, n, i, ki, i - 1);
    const int d = lp - lm >= 0 ? 1 : -1;
    const int dmin = d > 0 ? lm : lp;
    // The other end i + l d: l is the last m with lcp(i, i + m d) > dmin
    // (true up to l, false beyond).  The doubling probes 2, 4, ... are read
    // seven at a time (one round trip; most nodes need only the first
    // group); the last that holds, 2^q, bounds l in [2^q, 2^(q+1)).
    int lmax = 2;
    for (int q0 = 0;; q0 += 7) {
        uint64_t kp[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) kp[q] = key_at(k, n, i + (2 << (q0 + q)) * d);
        __builtin_amdgcn_sched_barrier(0);
        bool all = true;
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const bool tq = lcp_kb(n, i, ki, i + (2 << (q0 + q)) * d, kp[q]) > dmin;
            if (tq && all) lmax = 4 << (q0 + q);  // the probes that hold come first
            all = all && tq;
        }
        // (positions past the disk fail, so a disk's size bounds the groups)
        if (!all || q0 + 7 >= 28) break;
    }
    const int l = lmax / 2 + last_true(k, n, i, ki, d, lmax / 2, dmin, lmax / 2);
    const int jj = i + l * d;
    const int lo = min(i, jj), hi = max(i, jj);
    // the far end's key and the keys beside the range, read together
    const uint64_t kjj = key_at(k, n, jj), klo1 = key_at(k, n, lo - 1),
                   khi1 = key_at(k, n, hi + 1);
    __builtin_amdgcn_sched_barrier(0);
    const int dnode = lcp_kb(n, i, ki, jj, kjj);
    // split: the last m in [0, l) with lcp(i, i + m d) > dnode
    const int sp = last_true(k, n, i, ki, d, 0, dnode, l);
    const int gamma = i + sp * d + min(d, 0);
    brange[j] = make_int2(T.off[s] + lo, T.off[s] + hi);
    left[j] = lo == gamma ? ~(T.off[s] + gamma) : T.ioff[s] + gamma;
    right[j] = hi == gamma + 1 ? ~(T.off[s] + gamma + 1) : T.ioff[s] + gamma + 1;
    plen[j] = dnode;
    const int lev = oct_level(dnode);
    // parent prefix: max(lcp(lo - 1, lo), lcp(hi, hi + 1))
    const uint64_t klo = lo == i ? ki : kjj, khi = hi == i ? ki : kjj;
    const int dpar = max(lcp_kb(n, lo, klo, lo - 1, klo1), lcp_kb(n, hi, khi, hi + 1, khi1));
    octf[j] = (dpar < 0 || oct_level(dpar) != lev || lev == kLevels) ? 1 : 0;
    // cell centre: descend `lev` octant digits of this node's prefix
    const Root R = roots[s];
    float cx = R.c[0], cy = R.c[1], cz = R.c[2], hh = R.c[3];
    const uint64_t key = lo == i ? ki : kjj;  // k[lo], already read
    for (int q = 0; q < lev; ++q) {
        const int dg = (int)((key >> (3 * (kLevels - 1 - q))) & 7u);
        hh = hh * 0.5f;
        cx = cx + ((dg & 1) ? hh : -hh);
        cy = cy + ((dg & 2) ? hh : -hh);
        cz = cz + ((dg & 4) ? hh : -hh);
    }
    bin_cell[j] = make_float4(cx, cy, cz, hh);
}

// One launch for two independent passes over the sorted order: the first
// blocks build the binary radix tree, which never waits (dispatched first:
// its nodes with long ranges hold the longest chains of dependent reads);
// the next ntile blocks are the moment prefix (and the next ntile the
// second moments, with quadrupoles), each taking its tiles by ticket in
// dispatch order, so a tile waits only on tiles already running.
template <int KIND0>
__global__ __launch_bounds__(kBlock) void k_prefix_karras(
    SegTable T, const float4 *__restrict__ bodies, DVec<4> *P, DVec<4> *pagg, DVec<4> *pinc,
    DVec<6> *P2, DVec<6> *pagg2, DVec<6> *pinc2, int *tickets, int *ghist,
    const double *__restrict__ s64, const int *__restrict__ perm, int ntile,
    const uint64_t *__restrict__ keys, const Root *__restrict__ roots, int *left, int *right,
    int *plen, int *octf, float4 *bin_cell, int2 *brange)
{
    const int ntree = (T.ioff[T.nseg] + kBlock - 1) / kBlock;
    if ((int)blockIdx.x < ntree) {
        karras_node(T, keys, roots, left, right, plen, octf, bin_cell, brange,
                    (int)blockIdx.x * kBlock + (int)threadIdx.x);
        return;
    }
    const int b = (int)blockIdx.x - ntree;
    if (b < ntile) {
        prefix_tile<4, KIND0>(T, bodies, P, pagg, pinc, tickets, kTicketScan, ghist, s64, perm);
        return;
    }
    prefix_tile<6, 1>(T, bodies, P2, pagg2, pinc2, tickets, kTicketScan2, ghist, nullptr, nullptr);
}

// Mass and COM of sorted positions [q.x, q.y] of disk s from the moment
// prefix entries at the range's start (a) and past its end (b), divided once
// and rounded to float, and the traceless quadrupole about the COM when P2
// is given: Q = 3 C - tr(C) I with the central second moment C = S - M c c^T
// of the range sums S.  A massless range takes the midpoint of its end
// bodies.
__device__ inline float4 range_moments_of(const SegTable &T, int s, int2 q, const DVec<4> &a,
                                          const DVec<4> &b, const DVec<6> *__restrict__ P2,
                                          const float4 *__restrict__ bodies, float4 *quad)
{
    const int lo = q.x - T.off[s], hi = q.y - T.off[s];
    const double mm = b.v[0] - a.v[0];
    double cxd = 0, cyd = 0, czd = 0;
    float4 c;
    if (mm > 0.0) {
        cxd = (b.v[1] - a.v[1]) / mm;
        cyd = (b.v[2] - a.v[2]) / mm;
        czd = (b.v[3] - a.v[3]) / mm;
        c = make_float4((float)cxd, (float)cyd, (float)czd, (float)mm);
    } else {
        const float4 p = bodies[q.x], r = bodies[q.y];
        c = make_float4(0.5f * (p.x + r.x), 0.5f * (p.y + r.y), 0.5f * (p.z + r.z), 0.f);
    }
    if (P2) {
        const DVec<6> u = P2[pbase(T, s) + lo], o = P2[pbase(T, s) + hi + 1];
        const double sxx = o.v[0] - u.v[0] - mm * cxd * cxd;
        const double syy = o.v[1] - u.v[1] - mm * cyd * cyd;
        const double szz = o.v[2] - u.v[2] - mm * czd * czd;
        const double sxy = o.v[3] - u.v[3] - mm * cxd * cyd;
        const double sxz = o.v[4] - u.v[4] - mm * cxd * czd;
        const double syz = o.v[5] - u.v[5] - mm * cyd * czd;
        const double tr = sxx + syy + szz;
        quad[0] = make_float4((float)(3 * sxx - tr), (float)(3 * syy - tr), (float)(3 * szz - tr),
                              (float)(3 * sxy));
        quad[1] = make_float4((float)(3 * sxz), (float)(3 * syz), 0.f, 0.f);
    }
    return c;
}

// Enumerate the octree children of octree node j in Morton order: the
// binary descendants folded into it (inside one digit; bucket nodes of
// identical keys are never folded) are expanded level by level, each folded
// entry replaced in place by its two children, so the list stays in Morton
// order.  A digit leaves at most two folded levels below j's children, and
// each level reads its fold flags, then the children of its folded entries,
// in one round trip each.  Writes the codes to kids[0..n) and returns n, or
// -1 when the folds do not end within a digit (an inconsistent tree).
__device__ inline int oct_children(const int *__restrict__ left, const int *__restrict__ right,
                                   const int *__restrict__ octf, int j, int *kids)
{
    int cur[8];
    int n = 2;
    cur[0] = left[j];
    cur[1] = right[j];
#pragma unroll
    for (int lev = 0; lev < 3; ++lev) {
        // every flag and child is read at a clamped index whatever the
        // entry, so the loads of a level issue together
        int of[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) of[q] = octf[max(q < n ? cur[q] : 0, 0)];
        bool fold[8];
        int nf = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            fold[q] = q < n && cur[q] >= 0 && !of[q];
            nf += fold[q] ? 1 : 0;
        }
        if (nf == 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < n) kids[q] = cur[q];
            return n;
        }
        if (lev == 2 || n + nf > 8) return -1;
        int lc[8], rc[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = max(q < n ? cur[q] : 0, 0);
            lc[q] = left[c];
            rc[q] = right[c];
        }
        // entry q moves to pos[q] (two slots when folded): a selection
        // network with static indices keeps the lists in registers
        int pos[8], m = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            pos[q] = m;
            m += q < n ? (fold[q] ? 2 : 1) : 0;
        }
        int nxt[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            int v = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (q < n && pos[q] == o) v = fold[q] ? lc[q] : cur[q];
                if (q < n && fold[q] && pos[q] + 1 == o) v = rc[q];
            }
            nxt[o] = v;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) cur[q] = nxt[q];
        n = m;
    }
    return -1;
}

// Child counts.  Every internal node owns an 8-record block of the disk's
// records (first[j] = noff[s] + 8 + 8 (j - ioff[s]); the root record sits at
// noff[s]), so no allocation crosses threads: an octree node's children fill
// the head of its block, 8-aligned (one 128-byte line per record array), and
// the blocks of folded nodes stay unused.  The children found (at most 8
// codes per octree node) are kept in kids[8 j ..] for the record pass.
__global__ __launch_bounds__(kBlock) void k_oct_count(SegTable T, const int *__restrict__ left,
                                                      const int *__restrict__ right,
                                                      const int *__restrict__ octf, int *cnt,
                                                      int *first, int *err,
                                                      int *__restrict__ kids)
{
    const int ni = T.ioff[T.nseg];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ni) return;
    int c = 0;
    if (octf[j]) {
        c = oct_children(left, right, octf, j, kids + 8 * j);
        if (c < 0 || c > 8) {
            atomicOr(err, kErrTree);
            c = 0;
        }
    }
    cnt[j] = c;
    if (c) {
        const int s = seg_of_internal(T, j);
        first[j] = T.noff[s] + 8 + 8 * (j - T.ioff[s]);
    }
}

// Record count (root included) and deepest octree level of every disk, for
// the statistics call: block sums in LDS, one atomic per block and disk.
__global__ __launch_bounds__(kBlock) void k_tree_stats(SegTable T, const int *__restrict__ cnt,
                                                       const int *__restrict__ plen, int *nodes,
                                                       int *maxlev)
{
    __shared__ int s_tot[kMaxSeg], s_lev[kMaxSeg];
    if (threadIdx.x < kMaxSeg) {
        s_tot[threadIdx.x] = 0;
        s_lev[threadIdx.x] = 0;
    }
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < T.ioff[T.nseg] && cnt[j] > 0) {
        const int s = seg_of_internal(T, j);
        atomicAdd(&s_tot[s], cnt[j]);
        atomicMax(&s_lev[s], oct_level(plen[j]));
    }
    __syncthreads();
    if (threadIdx.x < T.nseg) {
        if (s_tot[threadIdx.x]) atomicAdd(&nodes[threadIdx.x], s_tot[threadIdx.x]);
        if (s_lev[threadIdx.x]) atomicMax(&maxlev[threadIdx.x], s_lev[threadIdx.x]);
    }
}

// Walk records: the children of octree node j into its block, the root
// record of each disk at noff[s], and one-body disks (root = the body).
// A thread takes kFillSlots child slots of one node and reads in three
// rounds for all of them at once: the node's count and child codes; each
// child's cell, range and count (a leaf: its body); the prefix entries at
// the ends of each child's range.  Its records are then stored together.
constexpr int kFillSlots = 2;

__global__ __launch_bounds__(kBlock) void k_oct_fill(
    SegTable T, const int *__restrict__ cnt, const DVec<4> *__restrict__ P,
    const DVec<6> *__restrict__ P2, const float4 *__restrict__ bin_cell,
    const float4 *__restrict__ bodies, float4 *__restrict__ rec_a, int4 *__restrict__ rec_b,
    float4 *__restrict__ rec_c, float4 *__restrict__ rec_q, const int2 *__restrict__ brange,
    int2 *__restrict__ rec_r, int *__restrict__ opar, int *__restrict__ opar_leaf,
    const int *__restrict__ kids)
{
    constexpr int SL = kFillSlots, PER = 8 / SL;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = tid / PER, k0 = (tid % PER) * SL;
    const int ni = T.ioff[T.nseg];
    // the record of binary node (or ~sorted position) `code` of disk s at r
    auto put = [&](int s, int code, int r, const float4 &cell, int2 rg, int nc, const DVec<4> &pa,
                   const DVec<4> &pb, const float4 &body) {
        float4 a, c;
        int4 b;
        float4 q[2] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
        if (code < 0) {
            a = body;
            b = make_int4(code, __float_as_int(-1.f), ~code, 0);
            c = make_float4(a.x, a.y, a.z, 0.f);
        } else {
            c = cell;
            a = range_moments_of(T, s, rg, pa, pb, rec_q ? P2 : nullptr, bodies, q);
            const int fc = T.noff[s] + 8 + 8 * (code - T.ioff[s]);
            b = make_int4(code, __float_as_int(c.w * c.w), fc, nc);
        }
        rec_a[r] = a;
        rec_b[r] = b;
        rec_c[r] = c;
        if (rec_q) {
            rec_q[2 * r] = q[0];
            rec_q[2 * r + 1] = q[1];
        }
        if (rec_r) rec_r[r] = code < 0 ? make_int2(~code, ~code) : rg;
    };
    // a disk's root record: its node 0, or its only body
    if (k0 == 0 && j < T.nseg) {
        const bool one = T.off[j + 1] - T.off[j] == 1;
        const int code = one ? ~T.off[j] : T.ioff[j];
        const int ic = one ? 0 : code;
        const int2 rg = one ? make_int2(T.off[j], T.off[j]) : brange[ic];
        const int nc = one ? 0 : cnt[ic];
        const float4 cell = bin_cell[ic], body = bodies[T.off[j]];
        const int base = T.off[j] + j;  // the disk's first prefix entry
        const DVec<4> pa = P[base + rg.x - T.off[j]], pb = P[base + rg.y - T.off[j] + 1];
        if (one) opar_leaf[T.off[j]] = -1;
        else opar[code] = -1;
        put(j, code, T.noff[j], cell, rg, nc, pa, pb, body);
    }
    if (j >= ni) return;
    // first round trip: the node's count and its child codes
    const int nk = cnt[j];
    int code[SL];
#pragma unroll
    for (int u = 0; u < SL; ++u) code[u] = kids[8 * j + k0 + u];
    if (nk <= k0) return;
    const int s = seg_of_internal(T, j);
    const int f0 = T.noff[s] + 8 + 8 * (j - T.ioff[s]);
    const int n_s = T.off[s + 1] - T.off[s];
    // second round trip: every child's cell, range and count, or its body (slots past
    // the count read safe places and store nothing)
    float4 cell[SL], body[SL];
    int2 rg[SL];
    int nc[SL];
#pragma unroll
    for (int u = 0; u < SL; ++u) {
        const bool ok = k0 + u < nk;
        const int ic = ok && code[u] >= 0 ? code[u] : T.ioff[s];
        const int lp = ok && code[u] < 0 ? ~code[u] : T.off[s];
        cell[u] = bin_cell[ic];
        rg[u] = brange[ic];
        nc[u] = cnt[ic];
        body[u] = bodies[lp];
    }
    // third round trip: the prefix at both ends of every child's range
    DVec<4> pa[SL], pb[SL];
    const int base = T.off[s] + s;
#pragma unroll
    for (int u = 0; u < SL; ++u) {
        const int lo = min(max(rg[u].x - T.off[s], 0), n_s - 1);
        const int hi = min(max(rg[u].y - T.off[s], 0), n_s - 1);
        pa[u] = P[base + lo];
        pb[u] = P[base + hi + 1];
    }
#pragma unroll
    for (int u = 0; u < SL; ++u) {
        if (k0 + u >= nk) break;
        // octree parent (binary index) of every record, for bottom-up searches
        if (code[u] < 0) opar_leaf[~code[u]] = j;
        else opar[code[u]] = j;
        put(s, code[u], f0 + k0 + u, cell[u], rg[u], nc[u], pa[u], pb[u], body[u]);
    }
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
                            const float *az, float dt_half, float dt, Bufs &B, hipStream_t st,
                            const LptDev &lpt)
{
    hipLaunchKernelGGL(k_kick_drift_bbox, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z,
                       vx, vy, vz, ax, ay, az, dt_half, dt, B.partial, lpt);
}

void launch_kick_drag(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                      const float *az, const float *dv, const float *x, const float *z,
                      float dt_half, int has_drag, float decay, float star_mass, int n,
                      int *step_ctr, hipStream_t st)
{
    hipLaunchKernelGGL(k_kick_drag, NBX_GRID(n), vx, vy, vz, ax, ay, az, dv, x, z, dt_half,
                       has_drag, decay, star_mass, n, step_ctr);
}

void launch_drag_kick_drift(const SegTable &T, float *x, float *y, float *z, float *vx,
                            float *vy, float *vz, const float *ax, const float *ay,
                            const float *az, const float *dv, float dt_half, float dt,
                            int has_drag, float decay, float star_mass, int *step_ctr, Bufs &B,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_drag_kick_drift, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z,
                       vx, vy, vz, ax, ay, az, dv, dt_half, dt, has_drag, decay, star_mass,
                       step_ctr, B.partial);
}

void launch_kick_drift_bbox_c(const SegTable &T, float *x, float *y, float *z, float *vx,
                              float *vy, float *vz, const float *ax, const float *ay,
                              const float *az, float dt_half, float dt, Bufs &B, hipStream_t st,
                              const LptDev &lpt)
{
    hipLaunchKernelGGL(k_kick_drift_bbox_c, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z,
                       vx, vy, vz, ax, ay, az, dt_half, dt, B.lo, B.partial, lpt);
}

void launch_kick_drag_c(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                        const float *az, const float *dv, const float *x, const float *z,
                        float dt_half, int has_drag, double decay, float star_mass, int n,
                        int *step_ctr, Bufs &B, hipStream_t st)
{
    hipLaunchKernelGGL(k_kick_drag_c, NBX_GRID(n), vx, vy, vz, ax, ay, az, dv, x, z, dt_half,
                       has_drag, decay, star_mass, n, step_ctr, B.lo);
}

void launch_diagnostics(const float *x, const float *y, const float *z, const float *vx,
                        const float *vy, const float *vz, const float *m, const float *lo,
                        const double *s64, int n, double star_mass, double soft, int with_self,
                        Bufs &B, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(64), 0, st, B.diag, 0.0, 8);
    hipLaunchKernelGGL(k_diag, NBX_GRID(n), x, y, z, vx, vy, vz, m, lo, s64, n, star_mass, soft,
                       B.diag);
    if (with_self)
        hipLaunchKernelGGL(k_diag_self, NBX_GRID(n), x, y, z, m, lo, s64, n, soft, B.diag);
}

void launch_load64(const float *x, const float *y, const float *z, const float *vx,
                   const float *vy, const float *vz, double *s64, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_load64, NBX_GRID(n), x, y, z, vx, vy, vz, s64, n);
}

void launch_state64(const float *x, const float *y, const float *z, const float *vx,
                    const float *vy, const float *vz, const float *ax, const float *ay,
                    const float *az, const float *lo, const double *s64, int n, double *out,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_state64, NBX_GRID(n), x, y, z, vx, vy, vz, ax, ay, az, lo, s64, n, out);
}

void launch_kick_drift_bbox_d(const SegTable &T, float *x, float *y, float *z, float *vx,
                              float *vy, float *vz, double dt_half, double dt, Bufs &B,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_kick_drift_bbox_d, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z,
                       vx, vy, vz, dt_half, dt, B.s64, B.partial);
}

void launch_kick_drag_d(float *vx, float *vy, float *vz, double dt_half, int has_drag,
                        double decay, double star_mass, int n, int *step_ctr, Bufs &B,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_kick_drag_d, NBX_GRID(n), vx, vy, vz, dt_half, has_drag, decay,
                       star_mass, n, step_ctr, B.s64);
}

void launch_bbox(const SegTable &T, const float *x, const float *y, const float *z, Bufs &B,
                 hipStream_t st)
{
    hipLaunchKernelGGL(k_bbox, dim3(T.blk[T.nseg]), dim3(kBlock), 0, st, T, x, y, z,
                       B.partial);
}

void launch_root(const SegTable &, Bufs &, hipStream_t) {}

template <int BITS>
static void launch_sort(const SegTable &T, const float *x, const float *y, const float *z,
                        const float *mass, Bufs &B, hipStream_t st)
{
    constexpr int R = 1 << BITS, NP = (63 + BITS - 1) / BITS;
    const int n = T.n_total;
    hipLaunchKernelGGL(k_keys<BITS>, dim3(cdiv(n, kKeysPerBlock)), dim3(kBlock), 0, st, T, x, y,
                       z, B.partial, B.roots, B.keys[0], B.vals[0], B.ghist, B.status[0],
                       R * B.n_sweep_tiles, (unsigned long long *)B.scanw,
                       kScanWords * B.n_scan_tiles, B.tickets);
    const int npass = NP + (T.nseg > 1 ? 1 : 0);
    int cur = 0;
    for (int p = 0; p < npass; ++p) {
        const bool last = p == npass - 1;
        const int mode = p < NP ? 0 : 1;
        hipLaunchKernelGGL(k_sweep<BITS>, dim3(B.n_sweep_tiles), dim3(kBlock), 0, st, B.keys[cur],
                           B.vals[cur], B.keys[cur ^ 1], B.vals[cur ^ 1], n, p, mode, T, B.ghist,
                           B.status[p & 1], last ? nullptr : B.status[(p + 1) & 1], B.tickets,
                           last ? 1 : 0, x, y, z, mass, B.bodies);
        cur ^= 1;
    }
    B.cur = cur;
}

// The wave-order placement on its own (builds without the fused key pass).
__global__ __launch_bounds__(kBlock) void k_lpt_scatter(LptDev L)
{
    lpt_scatter<kBlock>(L, blockIdx.x);
}

static void launch_lpt_scatter(const LptDev &lpt, hipStream_t st)
{
    if (lpt.nw > 0)
        hipLaunchKernelGGL(k_lpt_scatter, dim3(lpt_parts(lpt.nw)), dim3(kBlock), 0, st, lpt);
}

static void launch_split_sort(const SegTable &T, const float *x, const float *y, const float *z,
                              const float *mass, Bufs &B, bool resample, const LptDev &lpt,
                              hipStream_t st)
{
    const int n = T.n_total;
    if (resample || !B.split_ready) {
        launch_lpt_scatter(lpt, st);
        hipLaunchKernelGGL(k_keys<8>, dim3(cdiv(n, kKeysPerBlock)), dim3(kBlock), 0, st, T, x, y,
                           z, B.partial, B.roots, B.keys[0], nullptr, nullptr, nullptr, 0,
                           (unsigned long long *)B.scanw, kScanWords * B.n_scan_tiles,
                           B.tickets);
        hipLaunchKernelGGL(k_split, dim3(T.nseg), dim3(kSplitThreads), 0, st, T, B.keys[0],
                           B.spk, B.spi);
        hipLaunchKernelGGL(k_scatter, dim3(T.sblk[T.nseg]), dim3(kBlock), 0, st, T, B.keys[0], x,
                           y, z, mass, B.spk, B.spi, B.keys[1], B.vals[1], B.spay, B.bm);
    } else {
        hipLaunchKernelGGL(k_keys_scatter, dim3(T.sblk[T.nseg] + lpt_parts(lpt.nw)),
                           dim3(kBlock), 0, st, T, x, y, z, mass, B.partial, B.roots,
                           (unsigned long long *)B.scanw, kScanWords * B.n_scan_tiles, B.tickets,
                           B.spk, B.spi, B.keys[1], B.vals[1], B.spay, B.bm, lpt);
    }
    int nblk_max = 1;
    for (int q = 0; q < T.nseg; ++q) nblk_max = std::max(nblk_max, T.sblk[q + 1] - T.sblk[q]);
    hipLaunchKernelGGL(k_bucket_sort, dim3(T.bkoff[T.nseg]), dim3(kSortThreads),
                       sizeof(int) * (2 * nblk_max + 1), st, T, B.keys[1], B.vals[1], B.spay, B.bm,
                       B.keys[0], B.vals[0], x, y, z, mass, B.bodies, B.spk, B.spi, nblk_max);
    B.cur = 0;
    B.split_ready = true;
}

void launch_build(const SegTable &T, const float *x, const float *y, const float *z,
                  const float *mass, Bufs &B, hipStream_t st, bool resample, const LptDev &lpt,
                  const double *s64)
{
    const int n = T.n_total;
    if (B.radix_bits != 0) launch_lpt_scatter(lpt, st);
    switch (B.radix_bits) {
    case 0: launch_split_sort(T, x, y, z, mass, B, resample, lpt, st); break;
    case 9: launch_sort<9>(T, x, y, z, mass, B, st); break;
    case 11: launch_sort<11>(T, x, y, z, mass, B, st); break;
    default: launch_sort<8>(T, x, y, z, mass, B, st); break;
    }
    // the moment prefix(es) and the radix tree in one launch
    const int ni = T.ioff[T.nseg];
    const int ntile = B.n_scan_tiles;
    const int grid = (B.quad ? 2 * ntile : ntile) + cdiv(ni, kBlock);
    auto kern = s64 ? k_prefix_karras<2> : k_prefix_karras<0>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, T, B.bodies, B.prefix, B.pagg,
                       B.pinc, B.quad ? B.prefix2 : nullptr, B.pagg2, B.pinc2, B.tickets, B.ghist,
                       s64, B.vals[B.cur], ntile, B.keys[B.cur], B.roots, B.left, B.right,
                       B.plen, B.oct, B.bin_cell, B.brange);
}

void launch_tree_stats(const SegTable &T, Bufs &B, hipStream_t st)
{
    launch_fill_int(B.alloc, 0, 2 * kMaxSeg, st);
    const int ni = T.ioff[T.nseg];
    if (ni > 0)
        hipLaunchKernelGGL(k_tree_stats, NBX_GRID(ni), T, B.cnt, B.plen, B.alloc, B.maxlev);
}

void launch_summarize(const SegTable &T, Bufs &B, hipStream_t st)
{
    const int ni = T.ioff[T.nseg];
    if (ni > 0)
        hipLaunchKernelGGL(k_oct_count, NBX_GRID(ni), T, B.left, B.right, B.oct, B.cnt, B.first,
                           B.err, B.kids);
    hipLaunchKernelGGL(k_oct_fill, NBX_GRID(8 / kFillSlots * (ni > kMaxSeg ? ni : kMaxSeg)), T,
                       B.cnt, B.prefix, B.quad ? B.prefix2 : nullptr, B.bin_cell, B.bodies,
                       B.rec_a, B.rec_b, B.rec_c, B.quad ? B.rec_q : nullptr, B.brange, B.rec_r,
                       B.opar, B.opar_leaf, B.kids);
}

}  // namespace nbx
