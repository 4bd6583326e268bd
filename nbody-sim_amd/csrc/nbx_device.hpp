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

}  // namespace nbx
