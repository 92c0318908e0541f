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
// inner mean's error; 16x above the worst case of 2^-22-accurate normals and
// compensated sums.
constexpr double kMixedBound = 0x1p-16;
// FP64 redo of the mixed form: the first kRedoSlots listed outer samples with at
// most kRedoGroups final groups are split over kRedoParts blocks each.
constexpr uint32_t kRedoSlots = 65536, kRedoGroups = 64, kRedoParts = 32;  // model problem (see ModelPolicy::mixed_bound)

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
    uint32_t pad;
};

struct TestEntry {
    double s1, s2, sab;
    uint32_t item, done;
};

struct FinalEntry {
    double s1, sab;
    uint32_t item, group, done, kind;  // kind: 0 antithetic / fine-only, 1 plain fine, 2 plain coarse
};

// Device control block of one level call.
struct BatchedCtl {
    uint32_t cnt_f[kMaxRounds + 1], cnt_c[kMaxRounds + 1];
    uint32_t n_whole;  // outer samples whose final fits one segment (one lane, no atomics)
    uint32_t class_cnt[kMaxClasses], class_off[kMaxClasses + 1], class_cur[kMaxClasses];
    uint32_t final_total;
    uint32_t nreeval;
    uint32_t overflow;
    uint32_t pad;
    Partial acc;  // level sums, work and stats (atomics)
};

// Per-lane accumulation of one segment.
struct SegAcc {
    double s1 = 0., s2 = 0., sab = 0.;
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

// Per-lane level accumulators, flushed once per block.  They are touched once
// per finished segment or outer sample; the runtime-indexed arrays stay in
// scratch memory rather than taking registers from the inner loops.
struct LaneAcc {
    double f[kMaxPhi * kMaxMoments];
    double d[kMaxPhi * kMaxMoments];
    double work;
    uint64_t st[kStatsExt];
    __device__ void zero() {
        for (int i = 0; i < kMaxPhi * kMaxMoments; i++) f[i] = d[i] = 0.;
        work = 0.;
        for (int i = 0; i < kStatsExt; i++) st[i] = 0;
    }
};

// Adds every lane's accumulators into the level accumulator: a block-level
// reduction, then one atomic per field.  All threads of the block must call it.
__device__ __forceinline__ void block_flush(const LaneAcc& a, Partial* out) {
    __shared__ double sd[4];
    __shared__ uint64_t su[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int nsum = kMaxPhi * kMaxMoments;
    for (int k = 0; k < kPartialFields; k++) {
        const bool is_u = k >= 2 * nsum + 1;
        if (is_u) {
            uint64_t u = wave_sum_u64(a.st[k - (2 * nsum + 1)]);
            if (lane == 0) su[wave] = u;
        } else {
            double v = k < nsum ? a.f[k] : (k < 2 * nsum ? a.d[k - nsum] : a.work);
            v = wave_sum(v);
            if (lane == 0) sd[wave] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (is_u) {
                const uint64_t t = su[0] + su[1] + su[2] + su[3];
                if (t) atomicAdd((unsigned long long*)&out->stats[k - (2 * nsum + 1)], (unsigned long long)t);
            } else {
                const double t = ((sd[0] + sd[1]) + sd[2]) + sd[3];
                if (t != 0.) {
                    double* dst = k < nsum ? &out->f[k] : (k < 2 * nsum ? &out->d[k - nsum] : &out->work);
                    atomicAdd(dst, t);
                }
            }
        }
        __syncthreads();
    }
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
    uint32_t nps;        // normals per inner sample (draw diagnostics)
    uint32_t ups;        // uniforms per inner sample (draw diagnostics)
    uint32_t rounds;     // test rounds launched
    uint64_t base_min;   // smallest draw-group size: finals are base_min << class
    ItemState* items;
    TestEntry* tf[2];
    TestEntry* tc[2];
    FinalEntry* fpool;
    uint32_t fcap;
    uint32_t* reeval;
    double* redo_acc;    // [kRedoSlots x kRedoGroups] FP64 group sums of the redo
    uint32_t* redo_cnt;  // [kRedoSlots] parts of the redo done
    BatchedCtl* ctl;
    double* tr_phi_f;
    double* tr_phi_c;
    uint64_t* tr_nf;
    uint64_t* tr_nc;
};


// Defined once (reduce.hip): clears the control block; turns the final-group
// class counts into offsets and flags an over-full pool.
__global__ void k_b_zero(BatchedCtl* ctl);
__global__ void k_b_final_alloc(BatchedCtl* ctl, uint32_t fcap);

}  // namespace nmlmc
