// nbx_internal.hpp -- shared device/host definitions of the gfx950 disk step.
//
// Data layout in HBM (one nbx_tree, all disks of a batch back to back):
//   state      caller's ParticleData SoA in ORIGINAL body order (never
//              permuted, so id == index as in the reference, BHTree.cu:541)
//   keys/perm  63-bit Morton keys and body index, sorted (double-buffered
//              for the radix passes)
//   bodies     float4 {x,y,z,m} gathered in sorted order (walk leaves)
//   bin_*      binary radix tree over the sorted keys (n-1 nodes per disk)
//   rec_a/b    walk records of the compressed octree: float4 {com, mass}
//              and int4 {code, h^2, first child record, child count} (code:
//              the binary node of the record, ~sorted position for a leaf); the
//              children of internal node j fill the head of its own 8-record
//              block at noff[s] + 8 + 8 (j - ioff[s]); the root record of
//              disk s sits at noff[s].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbx {

constexpr int kMaxSeg = 16;      // disks per batched tree
constexpr int kLevels = 21;      // octant digits per key (63 bits)
constexpr int kBlock = 256;      // threads per block for streaming kernels
constexpr int kWave = 64;        // gfx950 wavefront
constexpr int kXcds = 8;         // MI355X accelerator dies; workgroups go to them round robin
constexpr int kStackCap = 384;   // gravity-walk stack entries per wave (LDS)
constexpr int kRecPad = 16;      // records of padding after the node array

// Splitter-bucket key sort (disks of at most kSplitMaxN bodies).
constexpr int kSplitItems = 2;                     // bodies per thread of a scatter block
constexpr int kSplitTile = kBlock * kSplitItems;   // bodies per scatter block
constexpr int kBucketMean = 320;                   // target bodies per bucket
constexpr int kMaxBuckets = 1024;                  // buckets per disk
constexpr int kOversample = 8;                     // samples per bucket
constexpr int kMaxSample = kOversample * kMaxBuckets;
constexpr int kBucketCap = 2048;                   // largest bucket sorted in LDS
constexpr int kSplitMaxN = 1 << 20;                // larger disks use the radix passes
constexpr int kMaxSplitBlocks = kSplitMaxN / kSplitTile;

// Per-disk offsets, passed by value to every kernel.
struct SegTable {
    int nseg;
    int n_total;                 // bodies over all disks
    int off[kMaxSeg + 1];        // body offset of disk s (prefix of counts)
    int blk[kMaxSeg + 1];        // kBlock-sized block offset of disk s
    int wave[kMaxSeg + 1];       // wave (64 targets) offset of disk s
    int ioff[kMaxSeg + 1];       // binary internal-node offset (n_s - 1 each)
    int noff[kMaxSeg + 1];       // walk-record offset (8 n_s + 8 each: root, 8 per node)
    int stile[kMaxSeg + 1];      // prefix-scan tile offset (tiles never straddle disks)
    int bkoff[kMaxSeg + 1];      // sort-bucket offset (kBucketMean bodies per bucket)
    int sblk[kMaxSeg + 1];       // scatter-block offset (kSplitTile bodies per block)
    int moff[kMaxSeg + 1];       // offset of the disk's (bucket x scatter block) count matrix
};

__host__ __device__ inline int seg_of_block(const SegTable &T, int b)
{
    int s = 0;
    while (s + 1 < T.nseg && b >= T.blk[s + 1]) ++s;
    return s;
}
__host__ __device__ inline int seg_of_body(const SegTable &T, int i)
{
    int s = 0;
    while (s + 1 < T.nseg && i >= T.off[s + 1]) ++s;
    return s;
}
__host__ __device__ inline int seg_of_wave(const SegTable &T, int w)
{
    int s = 0;
    while (s + 1 < T.nseg && w >= T.wave[s + 1]) ++s;
    return s;
}
__host__ __device__ inline int seg_of_stile(const SegTable &T, int t)
{
    int s = 0;
    while (s + 1 < T.nseg && t >= T.stile[s + 1]) ++s;
    return s;
}
__host__ __device__ inline int seg_of_bucket(const SegTable &T, int g)
{
    int s = 0;
    while (s + 1 < T.nseg && g >= T.bkoff[s + 1]) ++s;
    return s;
}
__host__ __device__ inline int seg_of_sblock(const SegTable &T, int b)
{
    int s = 0;
    while (s + 1 < T.nseg && b >= T.sblk[s + 1]) ++s;
    return s;
}
__host__ __device__ inline int seg_of_internal(const SegTable &T, int j)
{
    int s = 0;
    while (s + 1 < T.nseg && j >= T.ioff[s + 1]) ++s;
    return s;
}

// Wave-order sort of the next stepping walk (device view; the kernels that
// run its phases are listed with lpt_count / lpt_scatter).
constexpr int kLptBins = 1024;
constexpr int kLptPerBlock = 1024;  // waves placed by one scatter block

struct LptDev {
    const int *step_ctr = nullptr;
    const int *cost = nullptr;  // [2][stride]
    int *order = nullptr;       // [2][stride]
    int *hist = nullptr;        // [parts][kLptBins] cost histogram of every part
    int stride = 0, nw = 0;     // nw = 0: no sort
};

__host__ __device__ inline int lpt_parts(int nw) { return (nw + kLptPerBlock - 1) / kLptPerBlock; }

// Per-disk physical parameters of one step.
struct DiskDev {
    float r_coll;
    float pmass;                 // perturber mass in M_sun (0 = disabled)
    float px, py, pz;            // perturber position at the step's sim_time
};

struct DiskTable {
    DiskDev d[kMaxSeg];
};

// FoF link length per disk (passed by value).
struct LinkTable {
    float ll[kMaxSeg];
};

// Root cube per disk: {cx, cy, cz, half}.
struct Root {
    float c[4];
};

// Error flags written by kernels (host reads after the step).
enum ErrBits : int {
    kErrStack = 1,       // walk stack overflow
    kErrTree = 2,        // inconsistent tree (should not happen)
};

}  // namespace nbx
