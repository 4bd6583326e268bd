// nbx_fp64.hip -- FP64 node moments and FP64 tree walk (the FP64 step).
//
// The reference of the precision study: the same octree and the same
// acceptance rule as the float step (BHTree.cu:555-593: accept a node when
// h / r < theta, leaves always, skip the target's own leaf), with every
// distance, moment and force an FP64 operation on the FP64 state.  The cells
// (and so h) come from the keys of the rounded positions: they fix which
// bodies a node holds, not the precision of anything computed from them.
//
// One target per lane, targets in Morton order so a wave's paths stay close;
// a lane keeps one stack entry per level of its path in LDS, {first record,
// records left} of a child block, so the stack holds at most one entry per
// octree level plus the root: bounded by the 21 key levels.
#include <float.h>

#include "nbx_device.hpp"
#include "nbx_launch.hpp"

namespace nbx {

constexpr int kStack64 = kLevels + 3;  // root + one block per level + a level-21 bucket
constexpr int kDirect64 = 64;          // nodes of at most this many bodies sum their members

// FP64 {com, mass} of every walk record, indexed like the record pass: one
// thread per (internal node, child slot) plus the root of each disk.  A leaf
// is its body; a node of at most kDirect64 bodies sums its members directly;
// a larger node takes the difference of the FP64 prefix of {m, m x, m y, m z}
// summed from the FP64 positions (its cancellation error is at most ~1e-16 of
// the disk's moment over the node's mass, below 1e-12 AU here).
__device__ inline double4 moments64(const SegTable &T, int s, int code,
                                    const int2 *__restrict__ brange,
                                    const DVec<4> *__restrict__ P, const double *__restrict__ s64,
                                    const int *__restrict__ perm,
                                    const float4 *__restrict__ bodies)
{
    const size_t n = T.n_total;
    if (code < 0) {  // leaf: the body itself
        const int g = ~code, i = perm[g];
        return make_double4(s64[i], s64[n + i], s64[2 * n + i], (double)bodies[g].w);
    }
    const int2 q = brange[code];  // sorted positions [lo, hi] of the node
    double m = 0, cx = 0, cy = 0, cz = 0;
    if (q.y - q.x < kDirect64) {
        for (int g = q.x; g <= q.y; ++g) {
            const int i = perm[g];
            const double mm = bodies[g].w;
            m += mm;
            cx += mm * s64[i];
            cy += mm * s64[n + i];
            cz += mm * s64[2 * n + i];
        }
    } else {
        const int base = T.off[s] + s;  // the disk's first prefix entry
        const DVec<4> a = P[base + q.x - T.off[s]], b = P[base + q.y - T.off[s] + 1];
        m = b.v[0] - a.v[0];
        cx = b.v[1] - a.v[1];
        cy = b.v[2] - a.v[2];
        cz = b.v[3] - a.v[3];
    }
    if (m > 0.0) return make_double4(cx / m, cy / m, cz / m, m);
    // massless node: the midpoint of its end bodies (it pulls nothing)
    const int i = perm[q.x], j = perm[q.y];
    return make_double4(0.5 * (s64[i] + s64[j]), 0.5 * (s64[n + i] + s64[n + j]),
                        0.5 * (s64[2 * n + i] + s64[2 * n + j]), 0.0);
}

__global__ __launch_bounds__(kBlock) void k_rec64(SegTable T, const int *__restrict__ cnt,
                                                  const int *__restrict__ first,
                                                  const int *__restrict__ kids,
                                                  const int2 *__restrict__ brange,
                                                  const DVec<4> *__restrict__ P,
                                                  const double *__restrict__ s64,
                                                  const int *__restrict__ perm,
                                                  const float4 *__restrict__ bodies,
                                                  double4 *rec64)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = tid >> 3, k = tid & 7;
    if (k == 0 && j < T.nseg) {  // root of disk j: its node 0, or its one body
        const int code = T.off[j + 1] - T.off[j] == 1 ? ~T.off[j] : T.ioff[j];
        rec64[T.noff[j]] = moments64(T, j, code, brange, P, s64, perm, bodies);
    }
    if (j >= T.ioff[T.nseg] || k >= cnt[j]) return;
    const int s = seg_of_internal(T, j);
    rec64[first[j] + k] = moments64(T, s, kids[8 * j + k], brange, P, s64, perm, bodies);
}

__global__ __launch_bounds__(kBlock) void k_walk64(SegTable T, const double4 *__restrict__ rec64,
                                                   const int4 *__restrict__ rb,
                                                   const float4 *__restrict__ rc,
                                                   const int *__restrict__ perm, double *s64,
                                                   float *ax, float *ay, float *az, double theta2,
                                                   double eps2, double star_mass,
                                                   const DiskDev *__restrict__ dtab,
                                                   const int *__restrict__ step_ctr, int *err)
{
    // entry = first << 4 | records left (a child block holds at most 8)
    __shared__ int stk[kStack64][kBlock];
    const int tid = threadIdx.x;
    const int g = blockIdx.x * kBlock + tid;  // sorted position of the target
    if (g >= T.n_total) return;               // no barrier below
    const int s = seg_of_body(T, g);
    const size_t n = T.n_total;
    const int i = perm[g];
    const double X = s64[i], Y = s64[n + i], Z = s64[2 * n + i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    stk[0][tid] = T.noff[s] << 4 | 1;
    int sp = 1;
    while (sp > 0) {
        const int e = stk[sp - 1][tid];
        const int r = e >> 4;
        if ((e & 15) == 1) --sp;
        else stk[sp - 1][tid] = e + 15;  // next record, one left fewer
        const int4 b = rb[r];
        const double4 q = rec64[r];
        const double dx = q.x - X, dy = q.y - Y, dz = q.z - Z;
        const double rg2 = dx * dx + dy * dy + dz * dz;
        if (b.w > 0) {  // node: open unless h / r < theta
            const double h = rc[r].w;
            if (!(h * h < theta2 * rg2)) {
                if (sp == kStack64) {  // deeper than the key levels: cannot happen
                    atomicOr(err, kErrTree);
                    break;
                }
                stk[sp][tid] = b.z << 4 | b.w;
                ++sp;
                continue;
            }
        } else if (b.z == g) {
            continue;  // the target's own leaf
        }
        const double r2 = rg2 + eps2;
        const double f = q.w / (r2 * sqrt(r2));
        fx += f * dx;
        fy += f * dy;
        fz += f * dz;
    }
    // star at the origin (BHTree.cu:547-553) and the perturber's pull
    // (Perturber.cu:131-143), in FP64
    double ox, oy, oz;
    {
        const double dx = -X, dy = -Y, dz = -Z;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double f = star_mass / (r2 * sqrt(r2));
        ox = f * dx + fx;
        oy = f * dy + fy;
        oz = f * dz + fz;
    }
    DiskDev d{};
    if (dtab) d = dtab[*step_ctr * kMaxSeg + s];
    if (d.pmass > 0.f) {
        const double dx = (double)d.px - X, dy = (double)d.py - Y, dz = (double)d.pz - Z;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double f = (double)d.pmass / (r2 * sqrt(r2));
        ox += f * dx;
        oy += f * dy;
        oz += f * dz;
    }
    s64[6 * n + i] = ox;
    s64[7 * n + i] = oy;
    s64[8 * n + i] = oz;
    ax[i] = (float)ox;
    ay[i] = (float)oy;
    az[i] = (float)oz;
}

static inline int cdiv64(int a, int b) { return (a + b - 1) / b; }

void launch_rec64(const SegTable &T, Bufs &B, hipStream_t st)
{
    const int ni = T.ioff[T.nseg];
    hipLaunchKernelGGL(k_rec64, dim3(cdiv64(8 * (ni > kMaxSeg ? ni : kMaxSeg), kBlock)),
                       dim3(kBlock), 0, st, T, B.cnt, B.first, B.kids, B.brange, B.prefix, B.s64,
                       B.vals[B.cur], B.bodies, B.rec64);
}

void launch_walk64(const SegTable &T, const DiskDev *dtab, const int *step_ctr, double theta2,
                   double eps2, double star_mass, float *ax, float *ay, float *az, Bufs &B,
                   hipStream_t st)
{
    hipLaunchKernelGGL(k_walk64, dim3(cdiv64(T.n_total, kBlock)), dim3(kBlock), 0, st, T,
                       B.rec64, B.rec_b, B.rec_c, B.vals[B.cur], B.s64, ax, ay, az, theta2, eps2,
                       star_mass, dtab, step_ctr, B.err);
}

}  // namespace nbx
