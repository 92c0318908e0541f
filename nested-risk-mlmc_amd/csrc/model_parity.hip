// Plain FP64 form of the model problem with the reference's keying.
//
// Restates kernel_sample<false, model::sampler> (nested_sampler.hpp:369-485) for
// the OpenMP build's semantics: T virtual threads, thread t owning two streams
// keyed {seed, t} whose positions persist across calls, thread t taking the
// libgomp static-schedule block of [0, M) and walking its outer samples in
// order.  On the GPU a virtual thread is an aligned group of G lanes (G | 64),
// or a whole block of G = 256 lanes (four waves) when few virtual threads hold
// outer samples and their inner loops are long: the group walks the outer
// samples in order while its lanes split every inner loop, each lane computing
// the draws at the stream positions it owns.  Every decision (adaptive counts,
// phi, moments) is group-uniform, so the draws and the per-sample results
// equal those of the sequential reference.
#include "engine.hpp"
#include "model.hpp"

namespace nmlmc {

struct ModelParityArgs {
    ModelParams mp;
    EngineParams ep;
    LevelCounts cf, cc;      // fine (ell) and coarse (ell - 1) sample counts
    uint64_t seed;
    uint64_t M;
    uint32_t T;
    uint32_t ell;
    StreamState* states;     // [2 T]: sync, inner
    Partial* partials;       // [T]
    double* tr_phi_f;
    double* tr_phi_c;
    uint64_t* tr_nf;
    uint64_t* tr_nc;
};

// Sum over the G lanes of a virtual thread.  G = 256 is the whole block (one
// virtual thread per block, so every thread reaches the barriers): wave sums,
// then the four added in wave order.
template <int G>
__device__ __forceinline__ double vt_sum(double v) {
    if constexpr (G <= 64) {
        return group_sum<G>(v);
    } else {
        static_assert(G == 256, "a virtual thread is at most one 256-thread block");
        __shared__ double red[4];
        v = wave_sum(v);
        __syncthreads();  // the previous sum's reads are done
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        return ((red[0] + red[1]) + red[2]) + red[3];
    }
}

// Inner loop over `count` model samples from stream state st (the group's
// inner stream), lanes of the group splitting the Box-Muller pairs
// (inner_sampler::sample_inner, nested_sampler.hpp:131-161).
template <int VAR, int G>
__device__ __forceinline__ void parity_inner(const ModelParams& mp, uint64_t k0, uint64_t k1,
                                             StreamState& st, double Y, uint64_t count, bool want_s2,
                                             uint32_t gl, double& s1_out, double& s2_out) {
    constexpr int nps = (VAR == MODEL_REF) ? 2 : 1;
    const uint64_t K = count * nps;
    double s1 = 0., s2 = 0.;
    if (!st.is_saved && (st.P & 1) == 0) {
        const uint64_t pairs = (K + 1) >> 1;
        const uint64_t blk0 = st.P >> 1;
        for (uint64_t b = gl; b < pairs; b += G) {
            const u64x2 w = threefry2x64_20(blk0 + b, 0, k0, k1);
            double x, y;
            boxmuller_f64(w.a, w.b, x, y);
            if (VAR == MODEL_REF) {
                const double d = model_ref_inner(mp, Y, x, y);
                s1 += d;
                s2 += d * d;
            } else {
                const double d0 = Y + x;
                s1 += d0;
                s2 += d0 * d0;
                if (2 * b + 1 < count) {
                    const double d1 = Y + y;
                    s1 += d1;
                    s2 += d1 * d1;
                }
            }
        }
    } else {
        for (uint64_t i = gl; i < count; i += G) {
            double d;
            if (VAR == MODEL_REF) {
                d = model_ref_inner(mp, Y, normal_at_f64(k0, k1, st, 2 * i),
                                    normal_at_f64(k0, k1, st, 2 * i + 1));
            } else {
                d = Y + normal_at_f64(k0, k1, st, i);
            }
            s1 += d;
            s2 += d * d;
        }
    }
    s1_out = vt_sum<G>(s1);
    s2_out = want_s2 ? vt_sum<G>(s2) : 0.;
    st = advance_normals_f64(k0, k1, st, K);
}

// sample_inner_mean (nested_sampler.hpp:112-119).
template <int VAR, int G>
__device__ __forceinline__ double parity_mean(const ModelParams& mp, uint64_t k0, uint64_t k1,
                                              StreamState& st, double Y, uint64_t count, uint32_t gl) {
    const double start = -mp.max_loss;
    double s1, s2;
    parity_inner<VAR, G>(mp, k0, k1, st, Y, count, false, gl, s1, s2);
    return start + s1 / (double)count;
}

// sample_inner_delta2 (nested_sampler.hpp:120-129).
template <int VAR, int G>
__device__ __forceinline__ double parity_delta2(const ModelParams& mp, uint64_t k0, uint64_t k1,
                                                StreamState& st, double Y, uint64_t count,
                                                uint32_t gl) {
    const double start = -mp.max_loss;
    double s1, s2;
    parity_inner<VAR, G>(mp, k0, k1, st, Y, count, true, gl, s1, s2);
    const double out_var = s2 / (double)(count - 1) - (s1 * s1) / (double)(count * (count - 1));
    const double dist = start + s1 / (double)count;
    return dist * dist / out_var;
}

// get_samples_count::operator() (nested_sampler.hpp:240-273).
template <int VAR, int G>
__device__ __forceinline__ uint64_t parity_count(const ModelParams& mp, const EngineParams& ep,
                                                 const LevelCounts& lc, uint64_t k0, uint64_t k1,
                                                 StreamState& st, double Y, uint32_t gl,
                                                 uint64_t* stats, bool with_stats,
                                                 uint64_t& randn_count) {
    uint64_t cur = lc.base;
    constexpr int nps = (VAR == MODEL_REF) ? 2 : 1;
    while (true) {
        if (lc.max <= 2 * cur) {
            cur = lc.max;
            break;
        }
        stats[0] += cur;
        const double delta2 = parity_delta2<VAR, G>(mp, k0, k1, st, Y, cur, gl);
        randn_count += cur * nps;
        const double new_count =
            (lc.rule == 1) ? lc.tmp / delta2 : pow(delta2, -ep.adaptive_r / 2) * lc.tmp;
        if (new_count <= (double)cur) break;
        cur *= 2;
    }
    if (with_stats) {
        stats[1] += 1;
        stats[2] += (cur == lc.max);
        stats[3] += (cur == lc.base);
    }
    return cur;
}

// Waves per SIMD the kernel is compiled for: 2 keeps it at 255 VGPRs without
// spills (left alone the compiler reserves the whole register file and runs one
// wave a SIMD; the model sweep is 1.33x faster at 2; at 4 it spills 350-530 B
// a lane and is 1.18x).
#ifndef NMLMC_PLAIN_MODEL_WAVES
#define NMLMC_PLAIN_MODEL_WAVES 2
#endif

template <int VAR, int G>
__global__ __launch_bounds__(256, NMLMC_PLAIN_MODEL_WAVES) void k_model_parity(ModelParityArgs A) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t = gtid / G;
    const uint32_t gl = gtid % G;
    const bool active = t < A.T;  // whole groups share this flag (G divides the block)
    constexpr int nps = (VAR == MODEL_REF) ? 2 : 1;
    const ModelParams& mp = A.mp;
    const EngineParams& ep = A.ep;
    const uint64_t k0 = A.seed, k1 = t;

    // libgomp static schedule of [0, M) over T threads.
    uint64_t q = A.M / A.T, rr = A.M % A.T;
    if (t < rr) {
        q++;
        rr = 0;
    }
    const uint64_t s0 = active ? q * t + rr : 0, e0 = active ? s0 + q : 0;

    StreamState ss{}, si{};
    if (active) {
        ss = A.states[2 * t];
        si = A.states[2 * t + 1];
    }
    double f_sums[kMaxMoments] = {0, 0, 0, 0}, d_sums[kMaxMoments] = {0, 0, 0, 0};
    uint64_t stats[kStatsExt] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t randn_count = 0;
    const bool fine_only = (A.ell == 0);

    for (uint64_t m = s0; m < e0; m++) {
        // model::sampler::sample_outer (model_sampler.cpp:57-59).
        const double Y = st_randn_f64(k0, k1, ss);
        randn_count++;
        const uint64_t nf = parity_count<VAR, G>(mp, ep, A.cf, k0, k1, si, Y, gl, stats, true,
                                                 randn_count);
        uint64_t nc = 0;
        double f_phi = 0., c_phi = 0.;
        if (fine_only) {
            stats[0] += nf;
            f_phi = phi_scalar(mp.phi_type, parity_mean<VAR, G>(mp, k0, k1, si, Y, nf, gl), 1., 0.);
            randn_count += nf * nps;
        } else {
            nc = parity_count<VAR, G>(mp, ep, A.cc, k0, k1, si, Y, gl, stats, false, randn_count);
            if (ep.antithetic) {
                // sample_phi_antithetic (nested_sampler.hpp:171-216).
                const bool f_big = nf > nc;
                const uint64_t big = f_big ? nf : nc, small = f_big ? nc : nf;
                const uint64_t groups = big / small;
                const double coeff = (double)(1.f / (float)groups);
                const double start = -mp.max_loss;
                double s1_big = 0., phi_small = 0.;
                for (uint64_t g = 0; g < groups; g++) {
                    double s1, s2;
                    parity_inner<VAR, G>(mp, k0, k1, si, Y, small, false, gl, s1, s2);
                    phi_small = phi_scalar(mp.phi_type, start + s1 / (double)small, coeff, phi_small);
                    s1_big += s1;
                }
                const double phi_big = phi_scalar(mp.phi_type, start + s1_big / (double)big, 1., 0.);
                f_phi = f_big ? phi_big : phi_small;
                c_phi = f_big ? phi_small : phi_big;
                stats[0] += big;
                randn_count += big * nps;
            } else {
                stats[0] += nf + nc;
                f_phi = phi_scalar(mp.phi_type, parity_mean<VAR, G>(mp, k0, k1, si, Y, nf, gl), 1., 0.);
                c_phi = phi_scalar(mp.phi_type, parity_mean<VAR, G>(mp, k0, k1, si, Y, nc, gl), 1., 0.);
                randn_count += (nf + nc) * nps;
            }
        }
        add_moments(&f_phi, 1, f_sums, ep.fine_moments);
        if (!fine_only) {
            const double dphi = f_phi - c_phi;
            add_moments(&dphi, 1, d_sums, ep.diff_moments);
        }
        if (gl == 0) {
            if (A.tr_phi_f) A.tr_phi_f[m] = f_phi;
            if (A.tr_phi_c) A.tr_phi_c[m] = c_phi;
            if (A.tr_nf) A.tr_nf[m] = nf;
            if (A.tr_nc) A.tr_nc[m] = nc;
        }
    }
    if (active && gl == 0) {
        A.states[2 * t] = ss;
        A.states[2 * t + 1] = si;
    }
    // model::sampler::sample_inner adds one unit of work per inner sample drawn
    // (model_sampler.cpp:67), tests included: exactly the inner-sample count.
    const double work = (double)stats[0];
    stats[4] = 0;  // the model draws no uniforms
    stats[5] = randn_count;
    block_write_partial(active && gl == 0, f_sums, ep.fine_moments, d_sums, ep.diff_moments, work,
                        stats, &A.partials[blockIdx.x]);
}

template <int VAR>
static hipError_t launch_var(const ModelParityArgs& A, uint32_t G, uint32_t blocks, hipStream_t s) {
    const dim3 grid(blocks), block(256);
    switch (G) {
    case 1: hipLaunchKernelGGL((k_model_parity<VAR, 1>), grid, block, 0, s, A); break;
    case 2: hipLaunchKernelGGL((k_model_parity<VAR, 2>), grid, block, 0, s, A); break;
    case 4: hipLaunchKernelGGL((k_model_parity<VAR, 4>), grid, block, 0, s, A); break;
    case 8: hipLaunchKernelGGL((k_model_parity<VAR, 8>), grid, block, 0, s, A); break;
    case 16: hipLaunchKernelGGL((k_model_parity<VAR, 16>), grid, block, 0, s, A); break;
    case 32: hipLaunchKernelGGL((k_model_parity<VAR, 32>), grid, block, 0, s, A); break;
    case 256: hipLaunchKernelGGL((k_model_parity<VAR, 256>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((k_model_parity<VAR, 64>), grid, block, 0, s, A); break;
    }
    return hipGetLastError();
}

hipError_t launch_model_parity(const ModelParityArgsHost& H, hipStream_t s) {
    ModelParityArgs A;
    A.mp = H.mp;
    A.ep = H.ep;
    A.cf = H.cf;
    A.cc = H.cc;
    A.seed = H.seed;
    A.M = H.M;
    A.T = H.T;
    A.ell = H.ell;
    A.states = H.states;
    A.partials = H.partials;
    A.tr_phi_f = H.tr_phi_f;
    A.tr_phi_c = H.tr_phi_c;
    A.tr_nf = H.tr_nf;
    A.tr_nc = H.tr_nc;
    if (H.mp.variant == MODEL_REF) return launch_var<MODEL_REF>(A, H.G, H.blocks, s);
    return launch_var<MODEL_SHIFT>(A, H.G, H.blocks, s);
}

}  // namespace nmlmc
