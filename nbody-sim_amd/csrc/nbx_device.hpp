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

// Lanes 0..lane (2 << 63 wraps to 0, giving all 64 bits).
__device__ inline uint64_t lanemask_le() { return (2ull << __lane_id()) - 1ull; }

__device__ inline uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Inclusive add-scan over the 64 lanes of a wave: DPP row shifts by 1, 2, 4
// and 8 scan each row of 16 lanes, the broadcasts of lane 15 (to the row
// above, within each half) and of lane 31 (to the upper half) join the rows.
__device__ inline int wave_incl_add(int v)
{
    // Each term reads 0 where its source lane is outside the row (bound_ctrl)
    // or where its row is masked off (old = 0), so every step is one DPP add.
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 -> rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 -> rows 2, 3
    return v;
}

// Number of set bits of m below this lane (v_mbcnt).
__device__ inline int mbcnt64(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

}  // namespace nbx

// v_writelane_b32 (the intrinsic has no clang builtin here): lane `lane` of
// the result takes the wave-uniform `val`, every other lane keeps `old`.
extern "C" __device__ int nbx_writelane_i32(int val, int lane, int old) __asm(
    "llvm.amdgcn.writelane.i32");

namespace nbx {

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
// of a step reads order[p] and writes cost[p]; the sort, spread over the
// step's kernels, reads cost[p ^ 1] (the previous walk) and writes
// order[p ^ 1] (the next walk).  The waves are cut into parts of
// kLptPerBlock:
//   lpt_count    block `part` of the kick-drift pass writes the histogram of
//                its part's costs to hist[part][.] (LDS counters);
//   lpt_scatter  spare block `part` of the key pass derives its offsets from
//                the table (bin totals over all parts, counts of the earlier
//                parts) and places its waves with LDS counters.
// No global atomics; every histogram row is rewritten each step.
// ---------------------------------------------------------------------------

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

// Called by every block (kBlock threads) of the kick-drift pass.
__device__ inline void lpt_count(const LptDev &L)
{
    const int part = blockIdx.x;
    if (L.nw <= 0 || part >= lpt_parts(L.nw)) return;  // uniform per block
    constexpr int R = kLptPerBlock / kBlock;
    __shared__ int s_h[kLptBins];
    const int t = threadIdx.x;
    const int q = ((*L.step_ctr & 1) ^ 1) * L.stride;
    int bin[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int w = part * kLptPerBlock + r * kBlock + t;
        bin[r] = w < L.nw ? lpt_bin(L.cost[q + w]) : -1;
    }
    for (int b = t; b < kLptBins; b += kBlock) s_h[b] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (bin[r] >= 0) atomicAdd(&s_h[bin[r]], 1);
    __syncthreads();
    for (int b = t; b < kLptBins; b += kBlock) L.hist[part * kLptBins + b] = s_h[b];
}

// Called by all NT threads of the block (after every lpt_count has run).
template <int NT>
__device__ inline void lpt_scatter(const LptDev &L, int part)
{
    constexpr int BPT = kLptBins / NT, R = kLptPerBlock / NT;
    static_assert(kLptBins % NT == 0 && kLptPerBlock % NT == 0, "items per thread");
    __shared__ int s_off[kLptBins];
    __shared__ int s_w[NT / 64];
    const int t = threadIdx.x, lane = __lane_id(), wv = t / 64;
    const int nparts = lpt_parts(L.nw);
    const int q = ((*L.step_ctr & 1) ^ 1) * L.stride;
    int bin[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {  // this part's waves, loads in flight during the table sums
        const int w = part * kLptPerBlock + r * NT + t;
        bin[r] = w < L.nw ? lpt_bin(L.cost[q + w]) : -1;
    }
    int tot[BPT], bef[BPT];
#pragma unroll
    for (int u = 0; u < BPT; ++u) tot[u] = bef[u] = 0;
#pragma unroll 4
    for (int p = 0; p < nparts; ++p) {
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int c = L.hist[p * kLptBins + t * BPT + u];
            tot[u] += c;
            bef[u] += p < part ? c : 0;
        }
    }
    int sum = 0;
#pragma unroll
    for (int u = 0; u < BPT; ++u) sum += tot[u];
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
        s_off[t * BPT + u] = run + bef[u];
        run += tot[u];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (bin[r] >= 0) L.order[q + atomicAdd(&s_off[bin[r]], 1)] = part * kLptPerBlock + r * NT + t;
}

}  // namespace nbx
