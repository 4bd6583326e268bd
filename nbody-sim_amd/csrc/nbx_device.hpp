// nbx_device.hpp -- small device helpers shared by the gfx950 kernels.
#pragma once

#include "nbx_internal.hpp"

namespace nbx {

__device__ inline int lane_id() { return __lane_id(); }

__device__ inline uint64_t lanemask_lt()
{
    const int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }

__device__ inline int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-wide min / max through butterfly shuffles (64 lanes).
__device__ inline float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Block (kBlock threads) reduction of 3 mins and 3 maxes; thread 0 writes
// out[0..5] = {minx, miny, minz, maxx, maxy, maxz}.
__device__ inline void block_minmax6(float lx, float ly, float lz, float hx, float hy, float hz,
                                     float *out)
{
    __shared__ float s[kBlock / kWave][6];
    lx = wave_min(lx); ly = wave_min(ly); lz = wave_min(lz);
    hx = wave_max(hx); hy = wave_max(hy); hz = wave_max(hz);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        s[w][0] = lx; s[w][1] = ly; s[w][2] = lz;
        s[w][3] = hx; s[w][4] = hy; s[w][5] = hz;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = s[0][threadIdx.x];
        for (int k = 1; k < kBlock / kWave; ++k)
            v = threadIdx.x < 3 ? fminf(v, s[k][threadIdx.x]) : fmaxf(v, s[k][threadIdx.x]);
        out[threadIdx.x] = v;
    }
}

// Integrator.cu:48-75 for one body; unchanged when r_xz^2 < 1e-12.
__device__ inline void gas_drag_one(float xi, float zi, float &vxi, float &vyi, float &vzi,
                                    float decay, float star_mass)
{
#pragma clang fp contract(off)
    float r_xz2 = xi * xi + zi * zi;
    if (r_xz2 < 1e-12f) return;
    float r_xz = sqrtf(r_xz2);
    float inv_r = 1.0f / r_xz;
    float v_r = (vxi * xi + vzi * zi) * inv_r;
    float v_phi = (vxi * (-zi) + vzi * xi) * inv_r;
    float v_kep_mag = sqrtf(star_mass * inv_r);
    float v_kep = (v_phi >= 0.0f) ? v_kep_mag : -v_kep_mag;
    float v_r_new = v_r * decay;
    float v_phi_new = v_kep + (v_phi - v_kep) * decay;
    float v_y_new = vyi * decay;
    vxi = v_r_new * (xi * inv_r) + v_phi_new * (-zi * inv_r);
    vzi = v_r_new * (zi * inv_r) + v_phi_new * (xi * inv_r);
    vyi = v_y_new;
}

// Perturber.cu:131-143
__device__ inline void perturber_add_dev(float xi, float yi, float zi, float px, float py,
                                     float pz, float pmass, float soft, float &ax, float &ay,
                                     float &az)
{
#pragma clang fp contract(off)
    float dx = px - xi;
    float dy = py - yi;
    float dz = pz - zi;
    float r2 = dx * dx + dy * dy + dz * dz + soft * soft;
    float r = sqrtf(r2);
    float f = pmass / (r2 * r);
    ax += f * dx;
    ay += f * dy;
    az += f * dz;
}

// ---------------------------------------------------------------------------
// Wave order of the next stepping walk: a counting sort of the wave costs of
// the last walk, longest first (any permutation is correct; the order only
// shortens the tail of the launch).  Step parity p = *step_ctr & 1: the walk
// of a step reads order[p] and writes cost[p]; this sort, run during the
// tree build of the same step, reads cost[p ^ 1] (the previous walk) and
// writes order[p ^ 1] (the next walk).  One block of NT threads.
// ---------------------------------------------------------------------------
constexpr int kLptBins = 1024;
constexpr int kLptCache = 4096;  // costs held in registers by the block (the rest reloads)

// Monotone bin of a cost: exact below 64, then 32 bins per octave (< 864 in
// all).  Bins are listed in descending order, so the largest costs go first.
__device__ inline int lpt_bin(int cost)
{
    const int c = max(cost, 0);
    int b = c;
    if (c >= 64) {
        const int e = 31 - __clz(c);
        b = 64 + (e - 6) * 32 + ((c >> (e - 5)) & 31);
    }
    return kLptBins - 1 - b;
}

template <int NT>
__device__ inline void lpt_sort_block(const int *__restrict__ step_ctr,
                                      const int *__restrict__ cost, int *__restrict__ order,
                                      int stride, int nw)
{
    constexpr int NWV = NT / 64, BPT = kLptBins / NT, R = kLptCache / NT;
    static_assert(kLptBins % NT == 0 && kLptCache % NT == 0, "items per thread");
    __shared__ int s_h[kLptBins];
    __shared__ int s_w[NWV];
    const int t = threadIdx.x, lane = __lane_id(), wv = t / 64;
    const int q = ((*step_ctr & 1) ^ 1) * stride;
    const int *c = cost + q;
    int *o = order + q;
    int bn[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {  // one round of independent loads
        const int w = t + r * NT;
        bn[r] = w < nw ? lpt_bin(c[w]) : -1;
    }
    for (int b = t; b < kLptBins; b += NT) s_h[b] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (bn[r] >= 0) atomicAdd(&s_h[bn[r]], 1);
    // beyond the cache: rounds of 8 independent loads per thread
    for (int w0 = t + kLptCache; w0 < nw; w0 += 8 * NT) {
        int bb[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) bb[u] = w0 + u * NT < nw ? lpt_bin(c[w0 + u * NT]) : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (bb[u] >= 0) atomicAdd(&s_h[bb[u]], 1);
    }
    __syncthreads();
    // exclusive scan of the bins: BPT consecutive bins per thread
    int v[BPT], sum = 0;
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        v[u] = s_h[t * BPT + u];
        sum += v[u];
    }
    int inc = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    int run = inc - sum;
    for (int k = 0; k < wv; ++k) run += s_w[k];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        s_h[t * BPT + u] = run;
        run += v[u];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (bn[r] >= 0) o[atomicAdd(&s_h[bn[r]], 1)] = t + r * NT;
    for (int w0 = t + kLptCache; w0 < nw; w0 += 8 * NT) {
        int bb[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) bb[u] = w0 + u * NT < nw ? lpt_bin(c[w0 + u * NT]) : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (bb[u] >= 0) o[atomicAdd(&s_h[bb[u]], 1)] = w0 + u * NT;
    }
}

}  // namespace nbx
