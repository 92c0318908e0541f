// Batched engine: per-outer-sample keys, persistent grid-stride kernels over
// device work queues, one launch per refinement round.
//
// The reference walks each outer sample through get_samples_count (tests at
// N, 2N, ... discarded after each decision; nested_sampler.hpp:240-273) and
// then draws the final fine/coarse sums (171-216), using CUDA dynamic
// parallelism when one outer sample's inner loop is too large for a warp
// (316-351).  Here every draw group (one test round, or one antithetic small
// group of the final draws) is cut into segments of S inner units; any lane of
// the grid can take any segment because every draw is a pure function of
// (key, counter).  The lane that completes a draw group (a per-group counter,
// or a group of one segment) takes the decision at once and appends refined
// outer samples to the next round's queue at doubled count.  Finals are
// bucketed by size so each bucket is uniform; an outer sample is finalised by
// the lane that completes its last final group.
//
// Policy interface (see model_batched.hip, portfolio_batched.hip):
//   P::spu                      inner samples per unit
//   P::kTwoValued               final draws of levels > 0 give SegAcc::s1c too
//   P::init_item(A, m, item)    outer draw, start value (init_inner_sampler)
//   P::units<F32>(A, item_idx, item, tag, u0, u1, i0, i1, acc)
//                               accumulate units [u0, u1) of draw group
//                               [i0, i1) (samples) of draw group `tag`
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.hpp"

namespace nmlmc {

constexpr uint32_t TAG_FINE_TEST = 1, TAG_COARSE_TEST = 33, TAG_FINAL = 65, TAG_FINAL_COARSE = 66;
constexpr int kMaxRounds = 30;
constexpr int kMaxClasses = 32;

// Relative bound (times the mean magnitude of the inner terms) on the FP32
// inner mean's error.  FP32 normals lie within about 2^-20 of the FP64 ones
// (boxmuller_f32 takes the radius of the words whose float rounds to 1 from
// the complement word) and the compensated sums add 2^-23.
constexpr double kMixedBound = 0x1p-16;
// FP64 redo of the mixed form: the first kRedoSlots listed outer samples with at
// most kRedoGroups final groups are split over up to kRedoParts blocks each
// (one per 256 units of the level's largest final), in a grid of
// kRedoGridX x kRedoGridY blocks.  128 groups cover every antithetic final of
// the capped rule up to a cap of 2^6 (N_f / N_c <= 2^7); a level call lists at
// most a few thousand outer samples.
constexpr uint32_t kRedoSlots = 32768, kRedoGroups = 128, kRedoParts = 1024;
constexpr uint32_t kRedoGridX = 256, kRedoGridY = 32;

struct ItemState {
    double Y;        // model outer draw
    double start;    // init_inner_sampler value
    double phi_acc;  // sum over finished antithetic small groups of coeff * phi (atomic)
    double s1_big;   // (atomic)
    double sab_big;  // (atomic)
    double phi_f, phi_c;      // non-antithetic finals
    double phi32_f, phi32_c;  // mixed: FP32 phi values of a re-evaluated sample
    uint64_t nf, nc;
    uint32_t known;  // bit 0: fine count known, bit 1: coarse count known (atomic or)
    uint32_t left;   // unfinished final groups (atomic)
    uint32_t near;   // mixed: some final mean within the FP32 bound (atomic or)
    uint32_t inst;   // instance of a batch of problems (0 for a single problem)
};

struct TestEntry {
    double s1, s2, sab, smm;  // sab, smm: mixed form's error scales (SegAcc)
    uint32_t item, done;
};

struct FinalEntry {
    double s1, sab, s1c;  // s1c: coarse-discretisation sum (two-valued policies)
    uint32_t item, group, done, kind;  // kind: 0 antithetic / fine-only, 1 plain fine, 2 plain coarse
};

// Device control block of one level call.
struct BatchedCtl {
    uint32_t cnt_f[kMaxRounds + 1], cnt_c[kMaxRounds + 1];
    uint32_t n_whole;  // outer samples whose final fits one segment (one lane, no atomics)
    uint32_t class_cnt[kMaxClasses], class_off[kMaxClasses + 1], class_cur[kMaxClasses];
    uint32_t final_total;
    uint32_t nreeval;      // outer samples listed for the split FP64 redo (reeval[0 ..])
    uint32_t nreeval_seq;  // outer samples listed for the sequential redo (reeval[M - 1 ..] downwards)
    uint32_t overflow;
    uint32_t n_pend;  // fused first pass: outer samples left for the final phase
    uint32_t count_done;  // blocks of k_b_final_count that have counted
    uint32_t rc_f[kMaxRounds + 1], rc_c[kMaxRounds + 1];  // mixed: decisions of round r left to the FP64 recheck
    Partial acc;  // level sums, work and stats (atomics)
};

// Per-lane accumulation of one segment.
struct SegAcc {
    double s1 = 0., s2 = 0., sab = 0.;
    double smm = 0.;  // mixed: second-moment error scales of the FP32 terms (P::decision_bounds)
    // Two-valued policies (P::kTwoValued): the final draws' coarse values.  With
    // level-coupled inner paths an antithetic final inner sample has a value on
    // the level's discretisation (summed in s1) and one on the level below's,
    // from the same increments (summed here); the coarse estimator takes these.
    double s1c = 0.;
    double work = 0.;
    uint64_t nu = 0, nn = 0;  // uniforms and normals drawn (inner_stats[4], [5])
};

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// Per-lane level accumulators, flushed once per block.  The work and the
// draw statistics change once per segment; every index into st is a
// compile-time constant, so they live in registers (runtime-indexed arrays
// would go to scratch memory, and zeroing and flushing them over a grid of 2^19
// lanes costs ~70 us per launch whatever the work).  The moment sums change
// once per finished outer sample and live in LDS (lane_moments), so they take
// no registers from the inner loops.  Every kernel using them runs blocks of
// 256 threads.
struct LaneAcc {
    double work;
    uint64_t st[kStatsExt];
    __device__ __forceinline__ void zero() {
        work = 0.;
#pragma unroll
        for (int i = 0; i < kStatsExt; i++) st[i] = 0;
    }
};

// Moment sums of the block's lanes: the batched forms evaluate one scalar phi
// per outer sample, so kMaxMoments fine and kMaxMoments difference sums; sum
// j of lane t at [j * 256 + t] (fine j < kMaxMoments, then difference).
__device__ __forceinline__ double* lane_moments() {
    __shared__ double mom[2 * kMaxMoments * 256];
    return mom;
}
__device__ __forceinline__ void zero_lane_moments() {
    double* m = lane_moments() + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 2 * kMaxMoments; j++) m[j * 256] = 0.;
}

// add_moments for one scalar phi (nested_sampler.hpp:87-100): sum j += phi^(j+1)
// for j < moments, the powers formed by the same successive products; diff
// selects the difference sums.
__device__ __forceinline__ void add_lane_moments(double phi, uint32_t moments, bool diff) {
    double* m = lane_moments() + (diff ? kMaxMoments * 256 : 0) + threadIdx.x;
    double p = 1.;
#pragma unroll
    for (int j = 0; j < kMaxMoments; j++) {
        p *= phi;
        if ((uint32_t)j < moments) m[j * 256] += p;
    }
}

// Adds every lane's accumulators (and with MOM its moment sums) into the level
// accumulator: wave sums of all fields into LDS, one barrier, then one thread
// per field sums the four waves and issues one atomic.  All threads of the
// block must call it.
template <bool MOM>
__device__ __forceinline__ void block_flush(const LaneAcc& a, Partial* out) {
    constexpr int ND = 2 * kMaxMoments + 1;
    __shared__ double sd[ND][4];
    __shared__ uint64_t su[kStatsExt][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (MOM) {
        const double* m = lane_moments() + threadIdx.x;
#pragma unroll
        for (int k = 0; k < 2 * kMaxMoments; k++) {
            const double v = wave_sum(m[k * 256]);
            if (lane == 0) sd[k][wave] = v;
        }
    }
    {
        const double v = wave_sum(a.work);
        if (lane == 0) sd[ND - 1][wave] = v;
    }
#pragma unroll
    for (int k = 0; k < kStatsExt; k++) {
        const uint64_t v = wave_sum_u64(a.st[k]);
        if (lane == 0) su[k][wave] = v;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < ND && (MOM || t == ND - 1)) {
        const double s = ((sd[t][0] + sd[t][1]) + sd[t][2]) + sd[t][3];
        if (s != 0.) atomicAdd(t < kMaxMoments ? &out->f[t] : (t < 2 * kMaxMoments ? &out->d[t - kMaxMoments] : &out->work), s);
    } else if (t >= 64 && t < 64 + kStatsExt) {
        const int k = t - 64;
        const uint64_t s = su[k][0] + su[k][1] + su[k][2] + su[k][3];
        if (s) atomicAdd((unsigned long long*)&out->stats[k], (unsigned long long)s);
    }
}

// A grid-stride pass over work items [0, total) split into kXcds contiguous
// slices, slice x taken by the blocks of XCD x.  The dispatcher deals a grid's
// blocks out to the 8 XCDs round-robin (block b on XCD b % 8), so a plain
// grid stride sends neighbouring items to every XCD and each XCD's L2 ends up
// holding every item's data (outer states, item records); sliced, an XCD
// touches a disjoint eighth of the items.  The lanes of a wave still take
// consecutive items, active lanes forming a prefix of the wave.
constexpr uint32_t kXcds = 8;
struct XcdRange {
    uint64_t k0, end, stride;
};
__device__ __forceinline__ XcdRange xcd_range(uint64_t total, bool sliced = true) {
    const uint32_t g = gridDim.x, b = blockIdx.x, bd = blockDim.x;
    if (!sliced || g % kXcds) return XcdRange{(uint64_t)b * bd + threadIdx.x, total, (uint64_t)g * bd};
    const uint32_t x = b % kXcds, lb = b / kXcds;
    const uint64_t lo = total * x / kXcds, hi = total * (x + 1) / kXcds;
    return XcdRange{lo + (uint64_t)lb * bd + threadIdx.x, hi, (uint64_t)(g / kXcds) * bd};
}

// Cross-lane hand-off of partial sums.  Contributors add their sums with
// device-scope atomics, wait until those are acknowledged (s_waitcnt
// vmcnt(0): performed at device scope), then count their segment; the lane
// that completes the count reads the sums with atomic read-modify-writes.
// Nothing here goes through the non-coherent per-XCD L2 as a plain store, so
// no L2 write-back fence is needed.
__device__ __forceinline__ void aadd(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void astore(double* p, double v) {
    __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double aread(double* p) {
    return __hip_atomic_fetch_add(p, 0., __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t aread_u32(uint32_t* p) {
    return __hip_atomic_fetch_or(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wait_acks() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ uint32_t acount(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t aload_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Segmented sums over the lanes of a wave ahead of a cross-lane hand-off.  The
// active lanes holding segments of the same draw group (equal key; such lanes
// are adjacent, as consecutive lanes take consecutive segments) add their a
// and b; the last lane of each run gets the run's totals and its length n
// (tail = true).  One hand-off per run instead of one per segment: without it
// the ~1000 segments of one large group queue on the same few addresses.  All
// active lanes must call it together, and they must be a prefix of the wave
// (as in the grid-stride loops).
struct RunSums {
    double a, b;
    uint32_t n;
    bool tail;
};
__device__ __forceinline__ RunSums run_sums(unsigned long long key, double a, double b) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t active = __ballot(1);
    const unsigned long long kp = __shfl_up(key, 1u, 64);
    const uint64_t starts = __ballot(lane == 0 || kp != key);
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const double pa = __shfl_up(a, o, 64), pb = __shfl_up(b, o, 64);
        const unsigned long long pk = __shfl_up(key, o, 64);
        if (lane >= o && pk == key) {
            a += pa;
            b += pb;
        }
    }
    const bool tail = lane == 63 || !((active >> (lane + 1)) & 1) || ((starts >> (lane + 1)) & 1);
    const uint64_t at_or_below = starts & ((2ull << lane) - 1);  // lane 63: all bits
    const uint32_t first = 63 - (uint32_t)__clzll((long long)at_or_below);
    return {a, b, lane - first + 1, tail};
}

// Append with one atomic per wave for the lanes appending together.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const uint64_t mask = __ballot(want);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (mask) {
        const int first = __ffsll((long long)mask) - 1;
        if ((int)lane == first) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, first, 64);
    }
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    return base + (uint32_t)__popcll(mask & below);
}

// Fields shared by every problem's batched arguments (P::Args embeds it as .c).
struct BatchedCommon {
    EngineParams ep;
    LevelCounts cf, cc;
    uint64_t seed, M, offset;
    uint32_t ell;
    int phi_type;
    uint32_t S;          // inner units per segment
    uint32_t s_test_min; // smallest segment of a sparse later test round (0: rounds keep S)
    uint32_t redo_all;   // mixed form, tests: every Indicator outer sample takes the FP64 redo
    uint32_t nps;        // normals per inner sample (draw diagnostics)
    uint32_t ups;        // uniforms per inner sample (draw diagnostics)
    uint32_t rounds;     // test rounds launched
    uint64_t base_min;   // smallest draw-group size: finals are base_min << class
    ItemState* items;
    TestEntry* tf[2];
    TestEntry* tc[2];
    TestEntry* rq[2];    // mixed: fine / coarse test groups whose decision is rechecked in FP64
    FinalEntry* fpool;
    uint32_t fcap;
    uint32_t* reeval;
    uint32_t* pend;      // [M] outer samples the fused first pass leaves to the final phase
    uint32_t fused;      // 1: k_b_fused ran (round-0 tests in-lane; final phase over pend)
    double* redo_acc;    // [kRedoSlots x kRedoGroups] FP64 group sums of the redo
    uint32_t* redo_cnt;  // [kRedoSlots] parts of the redo done
    BatchedCtl* ctl;
    Partial* inst_out;   // batch of problems: per-instance level accumulators (else NULL: ctl->acc)
    uint32_t n_inst;     // instances in the batch (0 for a single problem)
    uint32_t zero_ctl;   // 1: the chain starts by clearing ctl (and inst_out); 0: ctl is clear already
    uint32_t xcd_sliced; // 1: grid-stride passes split per XCD (xcd_range)
    double* tr_phi_f;
    double* tr_phi_c;
    uint64_t* tr_nf;
    uint64_t* tr_nc;
};


// Level accumulators of a block, behind one interface for the two layouts:
//   MULTI = false (one problem): LaneAcc registers and the lanes' LDS moment
//     columns, flushed into c.ctl->acc;
//   MULTI = true (a batch of problems; every outer sample carries its
//     instance): per-instance sums in LDS, added with LDS atomics and flushed
//     into c.inst_out[instance].
// MOM: the kernel finishes outer samples (moment sums are kept).  init and
// flush must be reached by every thread of the block.
constexpr uint32_t kMaxInst = 64;
constexpr int kInstD = 2 * kMaxMoments + 1;  // fine moments, difference moments, work
__device__ __forceinline__ double* inst_sums() {
    __shared__ double a[kMaxInst * kInstD];
    return a;
}
__device__ __forceinline__ unsigned long long* inst_stats() {
    __shared__ unsigned long long a[kMaxInst * kStatsExt];
    return a;
}

template <bool MULTI>
struct BlockAcc {
    LaneAcc r;
    template <bool MOM>
    __device__ __forceinline__ void init(const BatchedCommon& c) {
        if (!MULTI) {
            r.zero();
            if (MOM) zero_lane_moments();
            return;
        }
        for (uint32_t i = threadIdx.x; i < c.n_inst * kInstD; i += blockDim.x) inst_sums()[i] = 0.;
        for (uint32_t i = threadIdx.x; i < c.n_inst * kStatsExt; i += blockDim.x) inst_stats()[i] = 0;
        __syncthreads();
    }
    __device__ __forceinline__ void work(uint32_t inst, double v) {
        if (MULTI) atomicAdd(&inst_sums()[inst * kInstD + 2 * kMaxMoments], v);
        else r.work += v;
    }
    __device__ __forceinline__ void st(uint32_t inst, int k, uint64_t v) {
        if (MULTI) {
            if (v) atomicAdd(&inst_stats()[inst * kStatsExt + k], (unsigned long long)v);
        } else {
            r.st[k] += v;
        }
    }
    // add_moments of phi (fine) and, when diff, of dphi (nested_sampler.hpp:87-100).
    __device__ __forceinline__ void moments(uint32_t inst, double phi, uint32_t fm, bool diff, double dphi,
                                            uint32_t dm) {
        if (!MULTI) {
            add_lane_moments(phi, fm, false);
            if (diff) add_lane_moments(dphi, dm, true);
            return;
        }
        double* a = inst_sums() + inst * kInstD;
        double p = 1.;
        for (uint32_t j = 0; j < fm; j++) {
            p *= phi;
            atomicAdd(&a[j], p);
        }
        if (diff) {
            p = 1.;
            for (uint32_t j = 0; j < dm; j++) {
                p *= dphi;
                atomicAdd(&a[kMaxMoments + j], p);
            }
        }
    }
    template <bool MOM>
    __device__ __forceinline__ void flush(const BatchedCommon& c) {
        if (!MULTI) {
            block_flush<MOM>(r, &c.ctl->acc);
            return;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < c.n_inst * kInstD; i += blockDim.x) {
            const double v = inst_sums()[i];
            const uint32_t q = i / kInstD, f = i - q * kInstD;
            Partial* out = &c.inst_out[q];
            if (v != 0.) atomicAdd(f < kMaxMoments ? &out->f[f] : (f < 2 * kMaxMoments ? &out->d[f - kMaxMoments] : &out->work), v);
        }
        for (uint32_t i = threadIdx.x; i < c.n_inst * kStatsExt; i += blockDim.x) {
            const unsigned long long v = inst_stats()[i];
            if (v) atomicAdd((unsigned long long*)&c.inst_out[i / kStatsExt].stats[i % kStatsExt], v);
        }
    }
};

// Defined once (reduce.hip): clears the control block (and a batch's
// per-instance accumulators).
__global__ void k_b_zero(BatchedCtl* ctl, Partial* inst_out, uint32_t n_inst);
// Launches k_b_zero on the control block alone.
hipError_t launch_clear_ctl(BatchedCtl* ctl, hipStream_t s);

// Turns the final-group class counts into offsets and flags an over-full pool;
// run by every thread of the block of k_b_final_count that counts last, after
// every block's counts were acknowledged.  The counts are read with atomics,
// one per thread side by side (read one after another they cost ~0.5 us each).
static_assert(kMaxClasses + 1 <= 256, "one thread per class count");
__device__ __forceinline__ void final_alloc(BatchedCtl* ctl, uint32_t fcap) {
    __shared__ uint32_t cnt[kMaxClasses + 1];
    const uint32_t t = threadIdx.x;
    if (t <= (uint32_t)kMaxClasses) cnt[t] = aread_u32(t < (uint32_t)kMaxClasses ? &ctl->class_cnt[t] : &ctl->n_whole);
    __syncthreads();
    if (t != 0) return;
    uint32_t off = 0;
    for (int i = 0; i < kMaxClasses; i++) {
        ctl->class_off[i] = off;
        ctl->class_cur[i] = 0;
        off += cnt[i];
    }
    ctl->class_off[kMaxClasses] = off;
    ctl->final_total = off + cnt[kMaxClasses];
    ctl->overflow = ctl->final_total > fcap;
    ctl->n_whole = 0;  // refilled as a cursor by k_b_final_fill
}

}  // namespace nmlmc
