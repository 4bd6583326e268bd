// nbx_wave.hpp -- wave-coherent octree walk shared by the force walk and its
// adjoint: a wavefront owns TG consecutive targets in Morton order and walks
// the compressed octree with a wave-uniform stack in LDS.  A stack entry is
// a block of child records {first, count} plus the mask of the targets whose
// own walk opened the parent.  Each iteration pops up to 64 records from the
// top of the stack, loads them with one lane-parallel vector load into an
// LDS staging area, and then evaluates them with LDS broadcasts; a target
// evaluates a record only if its bit is set, so every target visits exactly
// the nodes of its own per-body walk (BHTree.cu:555-593).
#pragma once

#include <type_traits>

#include "nbx_device.hpp"
#include "nbx_launch.hpp"

namespace nbx {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

constexpr int kWalkWaves = 1;   // one wavefront per block: walks never synchronise
constexpr int kWalkBlock = kWave * kWalkWaves;
// The gravity walk groups four independent wavefronts per block: a block's
// waves are placed one per SIMD of a compute unit, so every SIMD gets the
// same number of waves (single-wave blocks left some SIMDs with 3 and
// others with 5).  The waves never synchronise with each other.
constexpr int kGravWaves = 4;
constexpr int kGravBlock = kWave * kGravWaves;
constexpr int kGStack = kStackCap;  // gravity stack entries per wave
// Records of a popped batch are evaluated before their children are pushed,
// and the next batch is popped while the current one is evaluated, so up to
// two batches of 64 records (128 entries) can be pushed after a batched pop.
// Batches are taken only while that cannot pass the capacity; deeper stacks
// pop one entry at a time (depth-first), which adds at most 7 entries per
// octree level (plus the 8 children of the entry in flight), and a
// root-to-leaf path has at most 22 levels (21 digits and a bucket of
// identical keys, whose children are leaves).
constexpr int kDeepMargin = 7 * 22 + 8;
constexpr int kBatchMaxSp = kGStack - kDeepMargin - 2 * kWave;
static_assert(kBatchMaxSp >= 90, "stack too small for batched pops at theta = 0.5");

// A stack entry: child block {first, count} and the mask of the targets that
// opened its parent.  WIDE (64 targets per wave) keeps mask bits 32..63 in
// mhi; waves of at most 32 targets use 12-byte entries.
template <bool WIDE>
struct StackEntryT {
    int first, count;
    unsigned mlo, mhi;
};
template <>
struct StackEntryT<false> {
    int first, count;
    unsigned mlo;
};

// Staging slot of one popped record: {com, mass} and {mask bits 0..31, h^2,
// first child, child count} (the record's h is not used by the walks, its
// place carries the low half of the target mask).  Slots past the 64 of a
// batch absorb the discarded prefetch of a batch's last step.
struct StageSlot {
    float4 a;
    int4 b;
};
constexpr int kStageSlots = kWave + 8;

struct NoStageHi {};
struct StageHi {
    unsigned mhi[kStageSlots];  // mask bits 32..63 of the staged records
};

// LDS of one walking wave.  The footprint bounds the walk's occupancy (its
// registers allow 8 waves per SIMD): 7,168 bytes with at most 32 targets per
// wave (22 waves per CU), 8,992 bytes with 64.
template <bool WIDE>
struct WaveLdsT : std::conditional_t<WIDE, StageHi, NoStageHi> {
    static constexpr bool kWide = WIDE;
    StackEntryT<WIDE> st[kGStack];
    StageSlot s[kStageSlots];
    int own[kWave];             // pop: flags of the lanes that start a stack entry
};

template <bool WIDE>
struct WaveLdsQT : WaveLdsT<WIDE> {
    float4 q[2 * kStageSlots];
};

// Squared distance of the opening test with every rounding fixed: one
// product and two explicit fused multiply-adds, ((dx dx) + dy dy) + dz dz
// in that order, so the force walk, its interaction counts and its adjoint
// take identical decisions (the oracle's walk evaluates the same chain).
__device__ inline float dist2(float dx, float dy, float dz)
{
#pragma clang fp contract(off)
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Staging with the record index of every slot (visitors that write per
// record, such as the adjoint's node accumulators).
template <bool WIDE>
struct WaveLdsIdT : WaveLdsT<WIDE> {
    int id[kStageSlots];
};

// A visitor that declares `static constexpr bool kRecId = true` receives the
// record index in Rec::id.
template <class V, class = void>
struct wants_rec_id : std::false_type {};
template <class V>
struct wants_rec_id<V, std::void_t<decltype(V::kRecId)>> : std::integral_constant<bool, V::kRecId> {};

// Record data handed to a visitor (quadrupole rows only with QUAD).
struct Rec {
    float4 a;
    int4 b;
    float4 q0, q1;
    int id;
};

// Wave-level walk over TG targets.  Lane = r * TG + t serves target t and
// record slot r (L = 64 / TG slots): each popped batch of up to 64 records is
// evaluated L records at a time, lane (t, r) taking record k0 + r for target
// t.  Stack masks carry one bit per target.  With TG = 64 (L = 1) every lane
// owns one target and evaluates every record of the batch.
// Visitor: visit_masked(rec, mine) on every lane; mine says whether the
// lane's target owns the record.  The lane opens the (internal) record when
// mine is true and the returned value (accepted) is false.
// Returns the wave's work in record steps plus a per-batch overhead (the
// cost that orders the waves of the next launch, longest first).
template <int TG, bool QUAD, int U = 1, class Lds, class Visitor>
__device__ inline int wave_walk(const float4 *__restrict__ ra, const int4 *__restrict__ rb,
                                const float4 *__restrict__ rq, int root, uint64_t active, Lds &L,
                                int *err, Visitor &vis, int *wstat = nullptr)
{
    constexpr int LS = kWave / TG;
    constexpr uint64_t TMASK = TG == 64 ? ~0ull : ((1ull << TG) - 1ull);
    const int lane = __lane_id();
    const int tt = lane % TG, rr = lane / TG;
    // every lane writes the same entry: per-lane program order covers the
    // later reads, no cross-lane ordering is needed
    static_assert(TG <= 32 || Lds::kWide, "64 targets per wave need the wide stack entries");
    using Entry = std::remove_reference_t<decltype(L.st[0])>;
    auto entry = [](int first, int count, unsigned lo, unsigned hi) {
        Entry e;
        e.first = first;
        e.count = count;
        e.mlo = lo;
        if constexpr (Lds::kWide) e.mhi = hi;
        return e;
    };
    L.st[0] = entry(root, 1, (unsigned)active, (unsigned)(active >> 32));
    int sp = 1, sp_max = 1;
    int n_pop = 0, n_batch = 0, n_step = 0;  // records popped, batches, steps (wave-uniform)
    // Take up to 64 records from the top blocks (from the end of a block),
    // all lanes at once: lane j holds entry sp-1-j and the prefix sum of the
    // counts places entry j's records on lanes [excl_j, incl_j); the entry
    // crossing lane 64 keeps its first records on the stack.  A lane finds
    // the entry it serves from the flags of the entries' first lanes.
    auto pop = [&](int &nrec, int &my_rec, uint64_t &my_mask) {
        const bool deep = sp > kBatchMaxSp;
        const int ne = deep ? min(sp, 1) : min(sp, kWave);
        Entry top = entry(0, 0, 0u, 0u);
        if (lane < ne) top = L.st[sp - 1 - lane];
        const int incl = wave_incl_add(top.count);  // entries keep count >= 1
        const int excl = incl - top.count;
        nrec = min(__builtin_amdgcn_readlane(incl, kWave - 1), kWave);
        const int used = __popcll(ballot64(lane < ne && incl <= kWave));
        L.own[lane] = 0;
        if (lane < ne && excl < kWave) L.own[excl] = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint64_t starts = ballot64(L.own[lane] != 0);
        const int nst = (int)__popcll(starts & lanemask_le()) - 1;
        const int j = nst < 0 ? 0 : (nst > ne - 1 ? ne - 1 : nst);
        const int ej = __shfl(excl, j, kWave);
        my_rec = -1;
        my_mask = 0;
        if (lane < nrec) {
            const Entry e = L.st[sp - 1 - j];
            const int take = min(e.count, kWave - ej);
            const int off = min(max(e.count - take + (lane - ej), 0), e.count - 1);
            my_rec = e.first + off;
            my_mask = (uint64_t)e.mlo;
            if constexpr (Lds::kWide) my_mask |= (uint64_t)e.mhi << 32;
        }
        // (the partial entry is rewritten after every lane has read it)
        if (lane == used && lane < ne && excl < kWave) L.st[sp - 1 - lane].count = incl - kWave;
        sp -= used;
        n_pop += nrec;
        n_batch += nrec > 0;
    };
    // The record loads of the next batch are in flight (in registers) while
    // the current batch is evaluated: the next batch is popped from what the
    // stack holds before the current batch pushes its children.
    int nrec_n, rec_n;
    uint64_t mask_n;
    float4 ga = make_float4(0.f, 0.f, 0.f, 0.f), gq0 = ga, gq1 = ga;
    int4 gb = make_int4(0, 0, 0, 0);
    // Every lane loads: a lane without a record reads the disk's root record
    // (finite, staged under a zero mask), so the loads need no merge with the
    // previous values and stay in flight until the batch is staged.
    auto fetch = [&]() {
        pop(nrec_n, rec_n, mask_n);
        const int ri = rec_n >= 0 ? rec_n : root;
        // 32-bit byte offsets (records < 2^27) keep the loads on a scalar
        // base plus one vector offset
        const unsigned o16 = (unsigned)ri << 4;
        ga = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(ra) + o16);
        gb = *reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(rb) + o16);
        if constexpr (QUAD) {
            const char *q = reinterpret_cast<const char *>(rq) + 2 * o16;
            gq0 = *reinterpret_cast<const float4 *>(q);
            gq1 = *reinterpret_cast<const float4 *>(q + 16);
        }
    };
    fetch();
    while (nrec_n > 0) {
        const int nrec = nrec_n;
        // child block of the record this lane staged (pushed after the batch)
        const int rf = gb.z, rc = gb.w;
        {  // stage the batch that arrived; a lane without a record stages the
           // root record under a zero mask, so no slot needs a bound test
            L.s[lane].a = ga;
            // (h^2 staged times theta^-2: the opening test of every lane and
            // record is then one compare, h^2 theta^-2 < r^2)
            L.s[lane].b = make_int4((int)(unsigned)mask_n,
                                    __float_as_int(__int_as_float(gb.y) * vis.ith2), gb.z, gb.w);
            if constexpr (TG > 32) L.mhi[lane] = (unsigned)(mask_n >> 32);
            if constexpr (wants_rec_id<Visitor>::value) L.id[lane] = rec_n;
            if constexpr (QUAD) {
                L.q[2 * lane] = gq0;
                L.q[2 * lane + 1] = gq1;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        fetch();
        // Record steps, software-pipelined: the staging entries of the next
        // step are read while this one is evaluated (the staging area does
        // not change during the batch).
        auto stage = [&](int k, Rec &r, uint64_t &m) {
            r.a = L.s[k].a;
            r.b = L.s[k].b;
            m = (unsigned)r.b.x;
            if constexpr (TG > 32) m |= (uint64_t)L.mhi[k] << 32;
            if constexpr (wants_rec_id<Visitor>::value) r.id = L.id[k];
            if constexpr (QUAD) {
                r.q0 = L.q[2 * k];
                r.q1 = L.q[2 * k + 1];
            }
        };
        // a step covers SU = LS * U records: slot rr of unroll u takes record
        // k0 + u LS + rr, so every lane evaluates U independent records.
        // Steps run in pairs with one exit test per pair (both halves leave
        // their sums in the same registers, so the loop needs no copies).
        // Slots past the batch (below 64) are staged with zero masks, so the
        // odd step a pair may add evaluates nothing; the last prefetch reads
        // past the batch and is discarded.
        constexpr int SU = LS * U;
        static_assert(kWave % (2 * SU) == 0 && kWave + SU <= kStageSlots,
                      "pairs of steps tile the batch; the last prefetch stays inside the staging area");
        // Open masks of the batch: lane k of om_r (bits 32..63 in om_h)
        // receives the mask of the targets that opened record k.  Steps only
        // record them; the children are pushed once, after the batch.  With
        // at most 16 targets a 32-bit word packs RW = 32 / TG records' masks
        // (record k in lane k / RW, bits (k % RW) TG ..), so a step writes the
        // two halves of its ballot as they are.
        constexpr int RW = TG <= 16 ? 32 / TG : 1;
        int om_r = 0, om_h = 0;
        auto evaluate = [&](Rec(&cur)[U], const uint64_t(&mc)[U], int k0) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool mine = TG > 32 ? ((mc[u] >> tt) & 1ull) != 0
                                          : (((unsigned)mc[u] >> tt) & 1u) != 0;
                // (two ballots of plain compares: each is the compare's own
                // lane mask, with no select in between; the acceptance
                // compare serves both the pull and the open mask)
                const uint64_t om = ballot64(mine) & ~ballot64(vis.visit_masked(cur[u], mine));
                if constexpr (RW > 1) {
                    // records k0 + u LS .. (LS = 2 RW of them): two words
                    const int w = (k0 + u * LS) / RW;
                    om_r = nbx_writelane_i32((int)(unsigned)om, w, om_r);
                    om_r = nbx_writelane_i32((int)(unsigned)(om >> 32), w + 1, om_r);
                } else {
#pragma unroll
                    for (int q = 0; q < LS; ++q) {
                        // slot q of unroll u holds record k0 + u LS + q (< 64)
                        const int k = k0 + u * LS + q;
                        const uint64_t mq = (om >> (q * TG)) & TMASK;
                        om_r = nbx_writelane_i32((int)(unsigned)mq, k, om_r);
                        if constexpr (TG > 32)
                            om_h = nbx_writelane_i32((int)(unsigned)(mq >> 32), k, om_h);
                    }
                }
            }
        };
        // two register sets in turn (no copies between steps)
        Rec ca[U], cb[U];
        uint64_t ma[U], mb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) stage(u * LS + rr, ca[u], ma[u]);
        // (the scheduling barriers keep the next step's LDS reads ahead of
        // this step's arithmetic)
        for (int k0 = 0;;) {
#pragma unroll
            for (int u = 0; u < U; ++u) stage(k0 + SU + u * LS + rr, cb[u], mb[u]);
            __builtin_amdgcn_sched_barrier(0);
            evaluate(ca, ma, k0);
            k0 += SU;
#pragma unroll
            for (int u = 0; u < U; ++u) stage(k0 + SU + u * LS + rr, ca[u], ma[u]);
            __builtin_amdgcn_sched_barrier(0);
            evaluate(cb, mb, k0);
            n_step += 2;
            k0 += SU;
            if (k0 >= nrec) break;
        }
        {  // push the child blocks of the opened records in record order
           // (record k above record k - 1), lane k writing record k's entry
            // (an entry always holds records: the pop relies on counts >= 1)
            if constexpr (RW > 1)  // lane k takes record k's mask out of its word
                om_r = (int)(((unsigned)__shfl(om_r, lane / RW, kWave) >> ((lane % RW) * TG)) &
                             (unsigned)TMASK);
            const bool opened = lane < nrec && rc > 0 && (om_r != 0 || om_h != 0);
            const uint64_t fl = ballot64(opened);
            if (fl) {
                const int np = __popcll(fl);
                const int at = sp + mbcnt64(fl);
                if (opened && at < kGStack)
                    L.st[at] = entry(rf, rc, (unsigned)om_r, (unsigned)om_h);
                if (sp + np > kGStack && lane == 0) atomicOr(err, kErrStack);
                sp = min(sp + np, kGStack);
                sp_max = max(sp_max, sp);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (nrec_n == 0 && sp > 0) fetch();  // the stack was empty when prefetching
    }
    if (wstat && lane == 0) {  // per-wave work: records, batches, deepest stack
        wstat[0] = n_pop;
        wstat[1] = n_batch;
        wstat[2] = sp_max;
    }
    return n_step + 4 * n_batch;
}

// Sum of a per-lane value over the LS record slots of target tt.
template <int TG>
__device__ inline float slot_sum(float v)
{
#pragma unroll
    for (int o = TG; o < kWave; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <int TG>
__device__ inline int slot_sum_i(int v)
{
#pragma unroll
    for (int o = TG; o < kWave; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Gravity: accept a record when it is a leaf or h^2 theta^-2 < r^2 (the
// half-width / COM-distance test of BHTree.cu:576-577 without the square
// root and the division; the walk stages h^2 theta^-2, theta^-2 = 1 / (theta
// theta) rounded once, so a power-of-two theta scales exactly); skip the target's
// own leaf (BHTree.cu:580); monopole pull m d / (r^2+eps^2)^{3/2}.
// SELF = false drops the own-leaf test: a leaf record holds its body's
// position bit for bit, so for the own leaf d = +0 and the pull m d rinv^3
// is +0 whenever m rinv^3 is finite, which the launcher ensures by taking
// this form only for eps^2 >= kNoSelfTestEps2.
// ---------------------------------------------------------------------------
constexpr float kNoSelfTestEps2 = 1e-12f;

template <bool QUAD, bool SELF = true>
struct GravityVisitor {
    float bx, by, bz, ith2, eps2;  // ith2 = theta^-2 (the walk stages h^2 ith2)
    int self;
    float fx, fy, fz;
    // Branch-free: every lane evaluates the record and selects the result, so
    // the wave never splits its execution mask inside the record loop.
    __device__ inline bool visit_masked(const Rec &r, bool mine)
    {
        const float4 &a = r.a;
        const int4 &b = r.b;
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float r2 = dist2(dx, dy, dz);
        // a leaf stores h^2 = -1 (staged negative), so it is always accepted
        const bool acc = __int_as_float(b.y) < r2;
        bool take = mine & acc;
        if constexpr (SELF) take = take & !(b.w == 0 && b.z == self);
        const float rinv = __builtin_amdgcn_rsqf(r2 + eps2);
        const float rinv2 = rinv * rinv;
        const float rinv3 = rinv2 * rinv;
        // the pull is selected, not the mass: without softening the own
        // leaf has rinv = inf and m rinv^3 would be 0 inf = NaN
        // (with softening rinv is finite: the quadrupole form selects rinv^3
        // once and derives the pull and its r^-5, r^-7 factors from it)
        const float rsel = take ? rinv3 : 0.f;
        const float f = (QUAD && !SELF) ? a.w * rsel : (take ? a.w * rinv3 : 0.f);
        if constexpr (!QUAD) {
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
        } else {
            // a += -Q R / r^5 + (5/2) (R.Q R) R / r^7 for R = COM - body
            // (Q traceless; zero for leaves); the radial parts of the
            // monopole and the quadrupole share one multiply-add per axis
            const float qx = r.q0.x * dx + r.q0.w * dy + r.q1.x * dz;
            const float qy = r.q0.w * dx + r.q0.y * dy + r.q1.y * dz;
            const float qz = r.q1.x * dx + r.q1.y * dy + r.q0.z * dz;
            const float r5 = SELF ? (take ? rinv3 * rinv2 : 0.f) : rsel * rinv2;
            // (without softening the own leaf has rinv2 = inf: select, 0 inf = NaN)
            const float r7 = SELF ? (take ? r5 * rinv2 : 0.f) : r5 * rinv2;
            const float h = __builtin_fmaf(2.5f * r7, dx * qx + dy * qy + dz * qz, f);
            fx += h * dx;
            fy += h * dy;
            fz += h * dz;
            fx -= r5 * qx;
            fy -= r5 * qy;
            fz -= r5 * qz;
        }
        // the walk takes the open mask as ballot(mine) & ~ballot(acc)
        return acc;
    }
};

// Interaction counts of the same walk (accepted records excluding the own
// leaf, opened internal records): the algorithmic work of the force pass.
struct CountVisitor {
    float bx, by, bz, ith2;  // theta^-2 (the walk stages h^2 ith2)
    int self;
    int n_acc, n_open;
    __device__ inline bool visit_masked(const Rec &r, bool mine)
    {
        return !(mine && visit(r, r.b.w == 0));
    }
    __device__ inline bool visit(const Rec &r, bool leaf)
    {
        const float4 &a = r.a;
        const int4 &b = r.b;
        const float dx = a.x - bx, dy = a.y - by, dz = a.z - bz;
        const float r2 = dist2(dx, dy, dz);
        const bool open = !leaf && !(__int_as_float(b.y) < r2);
        if (open) ++n_open;
        else if (!(leaf && b.z == self)) ++n_acc;
        return open;
    }
};

// Waves per disk for a given number of targets per wave.
struct WaveMap {
    int wave[kMaxSeg + 1];
};

__device__ inline int seg_of_wmap(const SegTable &T, const WaveMap &M, int w)
{
    int s = 0;
    while (s + 1 < T.nseg && w >= M.wave[s + 1]) ++s;
    return s;
}

}  // namespace nbx
