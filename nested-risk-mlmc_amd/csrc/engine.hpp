// Shared definitions of the nested-sampler engine (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rng.hpp"

namespace nmlmc {

constexpr int kMaxPhi = 4;
constexpr int kMaxMoments = 4;
constexpr int kStats = 6;       // inner_stats reported through the ABI
constexpr int kStatsExt = 9;    // + mixed-form diagnostics: FP64 re-evaluations, flips, rechecked decisions

// One record of per-group partial results, reduced in a fixed order by
// k_reduce_partials: the per-level block of the reference (nested_sampler.hpp:587).
struct Partial {
    double f[kMaxPhi * kMaxMoments];
    double d[kMaxPhi * kMaxMoments];
    double work;
    uint64_t stats[kStatsExt];
};
constexpr int kPartialFields = 2 * kMaxPhi * kMaxMoments + 1 + kStatsExt;
// The reduced record of a level call is followed by this many staging records
// of the two-pass reduction (launch_reduce_partials).
constexpr uint32_t kReduceBlocks = 256;

// get_samples_count's constructor state (nested_sampler.hpp:225-239), computed on
// the host with the host libm so that tmp matches the reference bit for bit.
struct LevelCounts {
    uint64_t base;
    uint64_t max;
    double tmp;
    int adaptive;
    int rule;
};

struct EngineParams {
    uint32_t N0, beta;
    int antithetic;
    uint32_t fine_moments, diff_moments;
    double adaptive_r;
    uint32_t adapt_rule, cap_log2;
};

enum { PHI_INDICATOR = 0, PHI_MAX = 1, PHI_MOMENTS = 2, PHI_FMOMENTS = 3 };

// Sum over the G lanes of an aligned lane group with an xor butterfly: every
// lane of the group ends with the same, bit-identical, total.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <int G>
__device__ __forceinline__ uint64_t group_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of a double over the 64 lanes of a wave.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Lane `lane`'s value for every lane (lane the same in all lanes; it need not
// be active for the read, but its value is only meaningful when it is).
__device__ __forceinline__ uint32_t bcast_u32(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
__device__ __forceinline__ uint64_t bcast_u64(uint64_t v, uint32_t lane) {
    const uint32_t lo = bcast_u32((uint32_t)v, lane), hi = bcast_u32((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double bcast_f64(double v, uint32_t lane) {
    return __longlong_as_double((long long)bcast_u64((uint64_t)__double_as_longlong(v), lane));
}

// Reduces the group-uniform accumulators of a 256-thread block into one
// Partial record (only group leaders contribute).  Every thread of the block
// must call it.
__device__ __forceinline__ void block_write_partial(bool leader, const double* f, uint32_t nf,
                                                    const double* d, uint32_t nd, double work,
                                                    const uint64_t* st, Partial* out) {
    __shared__ double sd[4];
    __shared__ uint64_t su[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int k = 0; k < kPartialFields; k++) {
        const bool is_u = k >= 2 * kMaxPhi * kMaxMoments + 1;
        double v = 0.;
        uint64_t u = 0;
        if (leader) {
            if (k < kMaxPhi * kMaxMoments) v = (uint32_t)k < nf ? f[k] : 0.;
            else if (k < 2 * kMaxPhi * kMaxMoments) {
                const int j = k - kMaxPhi * kMaxMoments;
                v = (uint32_t)j < nd ? d[j] : 0.;
            } else if (!is_u) v = work;
            else u = st[k - (2 * kMaxPhi * kMaxMoments + 1)];
        }
        if (is_u) {
            u = wave_sum_u64(u);
            if (lane == 0) su[wave] = u;
        } else {
            v = wave_sum(v);
            if (lane == 0) sd[wave] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (is_u) out->stats[k - (2 * kMaxPhi * kMaxMoments + 1)] = su[0] + su[1] + su[2] + su[3];
            else {
                const double t = ((sd[0] + sd[1]) + sd[2]) + sd[3];
                if (k < kMaxPhi * kMaxMoments) out->f[k] = t;
                else if (k < 2 * kMaxPhi * kMaxMoments) out->d[k - kMaxPhi * kMaxMoments] = t;
                else out->work = t;
            }
        }
        __syncthreads();
    }
}

// add_moments (nested_sampler.hpp:87-100): out[i*m + j] += phi_i^(j+1).
__device__ __forceinline__ void add_moments(const double* phi, uint32_t phi_size, double* out,
                                            uint32_t moments) {
    for (uint32_t i = 0; i < phi_size; i++) {
        double p = 1.;
        for (uint32_t j = 0; j < moments; j++) {
            p *= phi[i];
            out[i * moments + j] += p;
        }
    }
}

// Single-component add_phi of the model sampler (model_sampler.cpp:81-91) and of
// the portfolio sampler's Indicator / Max cases (portfolio_sampler.cpp:261-291).
__device__ __forceinline__ double phi_scalar(int phi_type, double u, double coeff, double phi) {
    if (phi_type == PHI_MAX) return phi + coeff * (u > 0 ? u : 0.);
    return phi + coeff * (double)(u >= 0.);
}

}  // namespace nmlmc
