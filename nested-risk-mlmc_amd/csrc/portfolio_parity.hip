// Plain FP64 form of the portfolio problem with the reference's keying.
//
// kernel_sample<false, portfolio::sampler> (nested_sampler.hpp:369-485) for the
// OpenMP semantics: T virtual threads, thread t owning a sync and an inner
// stream keyed {seed, t} that persist across calls, its libgomp static share
// of [0, M) walked in order.  One wavefront per virtual thread:
//   - the outer draw (portfolio_sampler.cpp:293-311) takes its normals from
//     known positions of the sync stream, one stock per lane, into LDS;
//   - an inner loop with sub-sampling and no Unbiased options draws, for
//     sample k, the method uniform from sync word P_s + k and then 1 (BS) or 3
//     (Exact: option uniform + one Box-Muller pair) inner words; lanes take
//     chunks of 64 samples and find their inner word offsets by a ballot
//     prefix count, so every draw matches the sequential reference;
//   - other configurations (no sub-sampling, Unbiased options, a cached inner
//     normal) run the reference loop on thread 0.
// A virtual thread may take a block of W <= 8 waves (W > 1 only for books whose
// options are all Exact, chosen when few virtual threads have outer samples):
// the outer draw's stocks and the inner loop's samples are then spread over
// the block's lanes.  With every option Exact, sample k draws method 1 from
// sync word P_s + k whatever its value, and its option uniform and Box-Muller
// pair from inner words P_i + 3k .. P_i + 3k + 2, so any lane can take any
// sample; the lanes' sums are added wave by wave in a fixed order.
#include "portfolio.hpp"
#include "portfolio_dev.hpp"

namespace nmlmc {

struct PortfolioParityArgs {
    PortfolioDevice pf;
    EngineParams ep;
    LevelCounts cf, cc;
    uint64_t seed, M;
    uint32_t T, ell;
    double bs_work;
    StreamState* states;
    Partial* partials;  // one per virtual thread (block)
    double* tr_phi_f;
    double* tr_phi_c;
    uint64_t* tr_nf;
    uint64_t* tr_nc;
};

// Waves a virtual thread may take, and the doubles of LDS scratch a block
// needs beyond R_tau and z: the wave sums of a block of up to kMaxW waves (3
// each), or thread 0's broadcast (11 values).
constexpr uint32_t kMaxW = 8;
constexpr uint32_t kParityRed = 32;
static_assert(3 * kMaxW <= kParityRed && 11 <= kParityRed, "LDS reduction scratch");

struct InnerSums {
    double s1, s2, work;
    uint64_t nu, nn;
};

__device__ __forceinline__ double phi_value(int phi_type, double u, double coeff, double phi) {
    return phi_scalar(phi_type, u, coeff, phi);
}

// Kernel variants by the book's option methods: the general one; sub-sampled
// books of Exact options only (the only books given W > 1), whose kernel holds
// no BS or Unbiased code; books without Unbiased options.  Leaving code out
// keeps a kernel to a fraction of the general one's registers.
enum { PAR_GENERAL = 0, PAR_EXACT = 1, PAR_NO_UNBIASED = 2 };

// inner_sampler::sample_inner (nested_sampler.hpp:131-161) over `count` samples
// from the thread's streams (ss: sync, si: inner), block-uniform result.  red:
// LDS scratch of at least 3 W doubles.
template <uint32_t W, int KIND>
__device__ __forceinline__ InnerSums parity_inner(const PortfolioParityArgs& A, const double* R, uint64_t k1,
                                                  StreamState& ss, StreamState& si, uint64_t count, double* red) {
    constexpr bool EXACT_ONLY = KIND == PAR_EXACT;
    const PortfolioDevice& pf = A.pf;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    InnerSums r{0., 0., 0., 0, 0};
    const bool fast = pf.subsample != 0 && !(pf.compute_prob[2] > 0) && !si.is_saved;
    if (EXACT_ONLY && fast) {
        // Every option Exact: method 1 for every sample, three inner words each.
        double s1 = 0., s2 = 0., work = 0.;
        for (uint64_t k = tid; k < count; k += 64 * W) {
            const uint64_t q = si.P + 3 * k;
            const double u = u01_f64(stream_word(A.seed, k1, q));
            uint32_t idx;
            double prob;
            if (pf.subsample == -1) {
                idx = (uint32_t)(ceil(u * (double)pf.nopt[1]) - 1);
                prob = 1. / pf.options_count;
            } else {
                idx = pick_cdf(pf, 1, u * pf.compute_prob[1]);
                prob = pf.options[1][idx].prob / pf.prob_norm;
            }
            const OptionRec o = pf.options[1][idx];
            double x, y;
            stream_bm_f64(A.seed, k1, q + 1, x, y);
            const double v = inner_exact_d(pf, o, R, x, y, work) / prob;
            s1 += v;
            s2 += v * v;
        }
        const double a = wave_sum(s1), b = wave_sum(s2), c = wave_sum(work);
        if (lane == 0) {
            red[3 * wave] = a;
            red[3 * wave + 1] = b;
            red[3 * wave + 2] = c;
        }
        __syncthreads();
        double t1 = 0., t2 = 0., tw = 0.;
        for (uint32_t w = 0; w < W; w++) {
            t1 += red[3 * w];
            t2 += red[3 * w + 1];
            tw += red[3 * w + 2];
        }
        __syncthreads();  // red is reused by the next call
        ss.P += count;
        si.P += 3 * count;
        return InnerSums{t1, t2, tw, 2 * count, 2 * count};
    }
    if (!EXACT_ONLY && fast) {
        double s1 = 0., s2 = 0., work = 0.;
        uint64_t nn = 0;
        uint64_t base = 0;
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        for (uint64_t c0 = 0; c0 < count; c0 += 64) {
            const uint64_t k = c0 + lane;
            const bool valid = k < count;
            uint32_t method = 0;
            if (valid) method = pick_method(pf, u01_f64(stream_word(A.seed, k1, ss.P + k)) * pf.prob_norm);
            const uint64_t exact = __ballot(valid && method == 1);
            const uint64_t nvalid = count - c0 < 64 ? count - c0 : 64;
            const uint64_t q = si.P + base + (valid ? lane + 2 * __popcll(exact & below) : 0);
            if (valid) {
                const double u = u01_f64(stream_word(A.seed, k1, q));
                uint32_t idx;
                double prob;
                if (pf.subsample == -1) {
                    idx = (uint32_t)(ceil(u * (double)pf.nopt[method]) - 1);
                    prob = 1. / pf.options_count;
                } else {
                    idx = pick_cdf(pf, method, u * pf.compute_prob[method]);
                    prob = pf.options[method][idx].prob / pf.prob_norm;
                }
                const OptionRec o = pf.options[method][idx];
                double v;
                if (method == 1) {
                    double x, y;
                    stream_bm_f64(A.seed, k1, q + 1, x, y);
                    v = inner_exact_d(pf, o, R, x, y, work);
                    nn += 2;
                } else {
                    v = inner_bs_d(pf, o, R, work);
                }
                v = v / prob;
                s1 += v;
                s2 += v * v;
            }
            base += nvalid + 2 * (uint64_t)__popcll(exact);
        }
        r.s1 = wave_sum(s1);
        r.s2 = wave_sum(s2);
        r.work = wave_sum(work);
        r.nn = wave_sum_u64(nn);
        r.nu = 2 * count;
        ss.P += count;
        si.P += base;
        return r;
    }
    // Reference loop on thread 0.
    if (tid == 0) {
        DStream sy, in;
        sy.init(A.seed, k1, 0, ss.P, ss.saved, ss.is_saved);
        in.init(A.seed, k1, 0, si.P, si.saved, si.is_saved);
        double s1 = 0., s2 = 0., work = 0.;
        for (uint64_t i = 0; i < count; i++) {
            double d;
            if constexpr (EXACT_ONLY) {
                // pick_option and the Exact branch of inner_sample_d: the
                // method uniform is drawn, and method 1 is the only one.
                uint32_t method, idx;
                const double prob = pick_option(pf, sy, in, method, idx);
                const double z0 = in.randn();
                const double z1 = in.randn();
                d = inner_exact_d(pf, pf.options[1][idx], R, z0, z1, work) / prob;
            } else {
                d = inner_sample_d<KIND != PAR_NO_UNBIASED>(pf, R, sy, in, work);
            }
            s1 += d;
            s2 += d * d;
        }
        r = InnerSums{s1, s2, work, sy.nu + in.nu, sy.nn + in.nn};
        ss.P = sy.P;
        ss.saved = sy.saved;
        ss.is_saved = sy.is_saved;
        si.P = in.P;
        si.saved = in.saved;
        si.is_saved = in.is_saved;
    }
    if (W == 1) {
        r.s1 = __shfl(r.s1, 0, 64);
        r.s2 = __shfl(r.s2, 0, 64);
        r.work = __shfl(r.work, 0, 64);
        r.nu = __shfl(r.nu, 0, 64);
        r.nn = __shfl(r.nn, 0, 64);
        ss.P = __shfl(ss.P, 0, 64);
        ss.saved = __shfl(ss.saved, 0, 64);
        ss.is_saved = __shfl(ss.is_saved, 0, 64);
        si.P = __shfl(si.P, 0, 64);
        si.saved = __shfl(si.saved, 0, 64);
        si.is_saved = __shfl(si.is_saved, 0, 64);
        return r;
    }
    // W > 1: thread 0's results to the whole block through LDS.
    uint64_t* ru = reinterpret_cast<uint64_t*>(red);
    if (tid == 0) {
        red[0] = r.s1;
        red[1] = r.s2;
        red[2] = r.work;
        ru[3] = r.nu;
        ru[4] = r.nn;
        ru[5] = ss.P;
        red[6] = ss.saved;
        ru[7] = ss.is_saved;
        ru[8] = si.P;
        red[9] = si.saved;
        ru[10] = si.is_saved;
    }
    __syncthreads();
    r = InnerSums{red[0], red[1], red[2], ru[3], ru[4]};
    ss.P = ru[5];
    ss.saved = red[6];
    ss.is_saved = (uint32_t)ru[7];
    si.P = ru[8];
    si.saved = red[9];
    si.is_saved = (uint32_t)ru[10];
    __syncthreads();
    return r;
}

// Waves per SIMD the kernels are compiled for.  Exact-only, measured on the
// large and small books' plain sweeps: 2 (what the compiler picks unasked,
// 235-255 VGPRs, 336 B of scratch a lane) 18.5 and 13.2 ms, 3 (168 VGPRs,
// 624 B) 17.0 and 12.9 ms, 4 (128 VGPRs, 800 B) 17.4 and 13.0 ms.  The
// other two, left alone, get the whole register file and one wave a SIMD; at
// 2 a book of 4000 BS and 6000 Exact options sweeps (levels 0..4) in 24.1
// against 31.0 ms, one of 3000 BS, 5000 Exact and 2000 Unbiased options in
// 1.53 against 2.28 s (at 3 the spills make both slower).
#ifndef NMLMC_PLAIN_EXACT_WAVES
#define NMLMC_PLAIN_EXACT_WAVES 3
#endif
#ifndef NMLMC_PLAIN_NO_UNBIASED_WAVES
#define NMLMC_PLAIN_NO_UNBIASED_WAVES 2
#endif
#ifndef NMLMC_PLAIN_GENERAL_WAVES
#define NMLMC_PLAIN_GENERAL_WAVES 2
#endif

template <uint32_t W, int KIND>
__global__ __launch_bounds__(64 * W, (KIND == PAR_EXACT ? NMLMC_PLAIN_EXACT_WAVES : KIND == PAR_NO_UNBIASED ? NMLMC_PLAIN_NO_UNBIASED_WAVES : NMLMC_PLAIN_GENERAL_WAVES)) void k_portfolio_parity(
    PortfolioParityArgs A) {
    constexpr bool EXACT_ONLY = KIND == PAR_EXACT;
    extern __shared__ double smem[];  // R_tau[d], z[d], reduction scratch[kParityRed]
    const PortfolioDevice& pf = A.pf;
    const EngineParams& ep = A.ep;
    const uint32_t t = blockIdx.x, tid = threadIdx.x;
    const uint64_t k1 = t;
    const uint32_t d = pf.d;
    double* R = smem;
    double* z = smem + d;
    double* red = smem + 2 * d;

    uint64_t q = A.M / A.T, rr = A.M % A.T;
    if (t < rr) {
        q++;
        rr = 0;
    }
    const uint64_t s0 = q * t + rr, e0 = s0 + q;
    StreamState ss = A.states[2 * t], si = A.states[2 * t + 1];
    double f_sums[kMaxPhi * kMaxMoments] = {}, d_sums[kMaxPhi * kMaxMoments] = {};
    uint64_t stats[kStatsExt] = {0, 0, 0, 0, 0, 0, 0, 0};
    double work = 0.;
    const bool fine_only = A.ell == 0;
    const uint32_t ps = pf.phi_size;

    for (uint64_t m = s0; m < e0; m++) {
        // ---- portfolio::sampler::sample_outer ----
        uint64_t used;
        if (pf.chol) {
            for (uint32_t j = tid; j < d; j += 64 * W) z[j] = normal_at_f64(A.seed, k1, ss, j);
            __syncthreads();
            for (uint32_t i = tid; i < d; i += 64 * W) {
                double a = 0.;
                for (uint32_t j = 0; j <= i; j++) a += pf.chol[(size_t)i * d + j] * z[j];
                const StockRec st = pf.stocks[i];
                R[i] = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, pf.stau * a);
            }
            used = d;
        } else {
            const double dW_sys = pf.stau * normal_at_f64(A.seed, k1, ss, 0);
            if (d == 1) {
                if (tid == 0) R[0] = simulate_stock_d(pf.stocks[0].S0, pf.stocks[0].mu, pf.stocks[0].sigma, pf.tau, dW_sys);
                used = 1;
            } else {
                for (uint32_t i = tid; i < d; i += 64 * W) {
                    const double dW_ind = pf.stau * normal_at_f64(A.seed, k1, ss, 1 + i);
                    const double Wi = pf.rho * dW_sys + pf.s1_rho2 * dW_ind;
                    const StockRec st = pf.stocks[i];
                    R[i] = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, Wi);
                }
                used = 1 + d;
            }
        }
        __syncthreads();
        ss = advance_normals_f64(A.seed, k1, ss, used);
        stats[5] += used;
        // constant = -max_loss - compute_delta_cv() (sequential sum, as the reference).
        double cv = 0.;
        if (pf.use_delta_cv)
            for (uint32_t i = 0; i < d; i++) cv += (R[i] - pf.stocks[i].S0) * pf.delta_coalesce[i];
        const double constant = -pf.max_loss - cv;
        // init_inner_sampler: the pre-summed BS options without sub-sampling.
        double w_det = 0.;
        double start = constant;
        if constexpr (!EXACT_ONLY)
            if (pf.subsample == 0) start += det_bs_sum(pf, R, w_det);

        auto test_round = [&](uint64_t cur) {
            const InnerSums s = parity_inner<W, KIND>(A, R, k1, ss, si, cur, red);
            work += s.work + A.bs_work;
            stats[4] += s.nu;
            stats[5] += s.nn;
            const double out_var = s.s2 / (double)(cur - 1) - (s.s1 * s.s1) / (double)(cur * (cur - 1));
            const double dist = start + s.s1 / (double)cur;
            return dist * dist / out_var;
        };
        auto count_of = [&](const LevelCounts& lc, bool with_stats) {
            uint64_t cur = lc.base;
            while (true) {
                if (lc.max <= 2 * cur) {
                    cur = lc.max;
                    break;
                }
                stats[0] += cur;
                const double delta2 = test_round(cur);
                const double nc = (lc.rule == 1) ? lc.tmp / delta2 : pow(delta2, -ep.adaptive_r / 2) * lc.tmp;
                if (nc <= (double)cur) break;
                cur *= 2;
            }
            if (with_stats) {
                stats[1] += 1;
                stats[2] += (cur == lc.max);
                stats[3] += (cur == lc.base);
            }
            return cur;
        };
        auto mean_of = [&](uint64_t n) {
            const InnerSums s = parity_inner<W, KIND>(A, R, k1, ss, si, n, red);
            work += s.work + A.bs_work;
            stats[4] += s.nu;
            stats[5] += s.nn;
            return s.s1;
        };

        const uint64_t nf = count_of(A.cf, true);
        uint64_t nc = 0;
        double f_phi[kMaxPhi] = {}, c_phi[kMaxPhi] = {};
        if (fine_only) {
            stats[0] += nf;
            phi_add(pf, start + mean_of(nf) / (double)nf, 1., f_phi);
        } else {
            nc = count_of(A.cc, false);
            if (ep.antithetic) {
                const bool f_big = nf > nc;
                const uint64_t big = f_big ? nf : nc, small = f_big ? nc : nf;
                const uint64_t groups = big / small;
                const double coeff = (double)(1.f / (float)groups);
                double s1_big = 0.;
                double* out_small = f_big ? c_phi : f_phi;
                double* out_big = f_big ? f_phi : c_phi;
                for (uint64_t g = 0; g < groups; g++) {
                    const double s1 = mean_of(small);
                    phi_add(pf, start + s1 / (double)small, coeff, out_small);
                    s1_big += s1;
                }
                phi_add(pf, start + s1_big / (double)big, 1., out_big);
                stats[0] += big;
            } else {
                stats[0] += nf + nc;
                phi_add(pf, start + mean_of(nf) / (double)nf, 1., f_phi);
                phi_add(pf, start + mean_of(nc) / (double)nc, 1., c_phi);
            }
        }
        add_moments(f_phi, ps, f_sums, ep.fine_moments);
        if (!fine_only) {
            double dphi[kMaxPhi];
            for (uint32_t i = 0; i < ps; i++) dphi[i] = f_phi[i] - c_phi[i];
            add_moments(dphi, ps, d_sums, ep.diff_moments);
        }
        if (tid == 0) {
            if (A.tr_phi_f) A.tr_phi_f[m] = f_phi[0];
            if (A.tr_phi_c) A.tr_phi_c[m] = fine_only ? 0. : c_phi[0];
            if (A.tr_nf) A.tr_nf[m] = nf;
            if (A.tr_nc) A.tr_nc[m] = nc;
        }
        __syncthreads();
    }
    if (tid == 0) {
        A.states[2 * t] = ss;
        A.states[2 * t + 1] = si;
        Partial& p = A.partials[t];
        for (int i = 0; i < kMaxPhi * kMaxMoments; i++) {
            p.f[i] = f_sums[i];
            p.d[i] = d_sums[i];
        }
        p.work = work;
        for (int i = 0; i < kStatsExt; i++) p.stats[i] = stats[i];
    }
}

hipError_t launch_portfolio_parity(const PortfolioParityHost& H, hipStream_t s) {
    PortfolioParityArgs A;
    A.pf = H.pf;
    A.ep = H.ep;
    A.cf = H.cf;
    A.cc = H.cc;
    A.seed = H.seed;
    A.M = H.M;
    A.T = H.T;
    A.ell = H.ell;
    A.bs_work = H.bs_work;
    A.states = H.states;
    A.partials = H.partials;
    A.tr_phi_f = H.tr_phi_f;
    A.tr_phi_c = H.tr_phi_c;
    A.tr_nf = H.tr_nf;
    A.tr_nc = H.tr_nc;
    // Waves per virtual thread: more than one only for books whose options are
    // all Exact, while the virtual threads with outer samples (min(M, T)) are
    // too few to fill the device: up to kMaxW, doubling while they stay within
    // a budget of 6144 waves (the Exact-only kernel runs three waves a SIMD;
    // the large and small books' plain sweeps took 16.9 and 12.5 ms at a
    // budget of 1536, 16.3 and 11.7 ms at 6144, 16.5 and 11.5 ms at 12288).
    // NMLMC_PLAIN_WAVE_BUDGET and NMLMC_PLAIN_WAVES (1, 2, 4 or 8) override it;
    // per-sample results do not depend on it.
    uint32_t W = 1;
    const bool exact_only = H.pf.subsample != 0 && H.pf.nopt[0] == 0 && H.pf.nopt[2] == 0;
    if (exact_only) {
        const uint64_t act = H.M < (uint64_t)H.T ? H.M : (uint64_t)H.T;
        uint64_t budget = 6144;
        if (const char* env = getenv("NMLMC_PLAIN_WAVE_BUDGET")) budget = strtoull(env, nullptr, 10);
        uint32_t wmax = kMaxW;
        if (const char* env = getenv("NMLMC_PLAIN_WAVES_MAX")) wmax = (uint32_t)strtoul(env, nullptr, 10);
        while (W < wmax && W < kMaxW && act * (2 * W) <= budget) W *= 2;
        if (const char* env = getenv("NMLMC_PLAIN_WAVES")) {
            const uint32_t v = (uint32_t)strtoul(env, nullptr, 10);
            W = (v >= 8) ? 8 : (v >= 4) ? 4 : (v >= 2) ? 2 : 1;
        }
    }
    const size_t lds = sizeof(double) * (2 * (size_t)H.pf.d + kParityRed);
    switch (W) {
    case 8: hipLaunchKernelGGL((k_portfolio_parity<8, PAR_EXACT>), dim3(H.blocks), dim3(512), lds, s, A); break;
    case 4: hipLaunchKernelGGL((k_portfolio_parity<4, PAR_EXACT>), dim3(H.blocks), dim3(256), lds, s, A); break;
    case 2: hipLaunchKernelGGL((k_portfolio_parity<2, PAR_EXACT>), dim3(H.blocks), dim3(128), lds, s, A); break;
    default:
        if (exact_only)
            hipLaunchKernelGGL((k_portfolio_parity<1, PAR_EXACT>), dim3(H.blocks), dim3(64), lds, s, A);
        else if (H.pf.nopt[2] == 0)
            hipLaunchKernelGGL((k_portfolio_parity<1, PAR_NO_UNBIASED>), dim3(H.blocks), dim3(64), lds, s, A);
        else
            hipLaunchKernelGGL((k_portfolio_parity<1, PAR_GENERAL>), dim3(H.blocks), dim3(64), lds, s, A);
        break;
    }
    return hipGetLastError();
}

// sampler::sample_delta_unbiased (portfolio_sampler.cpp:397-444): pathwise
// delta of an Unbiased option from an antithetic pair, dt_h0inv Milstein steps
// (volatility sigma and -sigma) against the exact GBM paths driven by the same
// Brownian increment.
__device__ double delta_unbiased_d(const PortfolioDevice& pf, const OptionRec& o, DStream& in, double& work) {
    const StockRec st = pf.stocks[o.stock_ind];
    const double r = pf.r, sigma = st.sigma;
    const double T = o.maturity - pf.tau;
    const uint64_t N = pf.dt_h0inv;
    const double dW_0_tau = pf.stau * in.randn();
    const double S1 = simulate_stock_d(st.S0, r, sigma, pf.tau, dW_0_tau);
    const double S2 = simulate_stock_d(st.S0, r, sigma, pf.tau, -dW_0_tau);
    double STf[2] = {S1, S2};
    const double sig[2] = {sigma, -sigma};
    const double dt = T / (double)N;
    const double sqrt_dt = sqrt(dt);
    double W = 0.;
    for (uint64_t n = 0; n < N; n++) {
        const double dW = sqrt_dt * in.randn();
        W += dW;
        for (int i = 0; i < 2; i++) {
            const double m = pf.use_milstein ? (0.5 * (sig[i] * sig[i]) * (dW * dW - dt)) : 0.;
            STf[i] += STf[i] * (r * dt + sig[i] * dW + m);
        }
    }
    const double ST0 = simulate_stock_d(S1, r, sigma, T, W);
    const double ST1 = simulate_stock_d(S2, r, sigma, T, -W);
    work += (double)(2 * N) * o.W_est;
    double val = 0.;
    val += derv_from_stock_d(o, STf[0], st.S0, r);
    val += derv_from_stock_d(o, STf[1], st.S0, r);
    val -= derv_from_stock_d(o, ST0, st.S0, r);
    val -= derv_from_stock_d(o, ST1, st.S0, r);
    return val / (2.0 * o.w);
}

// One lane per virtual thread t: its libgomp static share of the M samples, in
// order, from its inner stream {seed, t}; moments and work go to partials[t].
// The work is tiny (dt_h0inv steps per sample), so lanes are independent and no
// lane of a block waits on another.
__global__ __launch_bounds__(64) void k_portfolio_delta(PortfolioDeltaHost A) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.T) return;
    uint64_t q = A.M / A.T, rr = A.M % A.T;
    if (t < rr) {
        q++;
        rr = 0;
    }
    const uint64_t s0 = q * t + rr, e0 = s0 + q;
    StreamState si = A.states[2 * t + 1];
    DStream in;
    in.init(A.seed, t, 0, si.P, si.saved, si.is_saved);
    const OptionRec o = A.pf.options[2][A.option_index];
    double f[kMaxMoments] = {0., 0., 0., 0.};
    double work = 0.;
    for (uint64_t m = s0; m < e0; m++) {
        const double dl = delta_unbiased_d(A.pf, o, in, work);
        double p = 1.;
        for (uint32_t j = 0; j < A.moments; j++) {
            p *= dl;
            f[j] += p;
        }
    }
    si.P = in.P;
    si.saved = in.saved;
    si.is_saved = in.is_saved;
    A.states[2 * t + 1] = si;
    Partial& P = A.partials[t];
    for (int i = 0; i < kMaxPhi * kMaxMoments; i++) {
        P.f[i] = i < kMaxMoments ? f[i] : 0.;
        P.d[i] = 0.;
    }
    P.work = work;
    for (int i = 0; i < kStatsExt; i++) P.stats[i] = 0;
}

hipError_t launch_portfolio_delta(const PortfolioDeltaHost& H, hipStream_t s) {
    const uint32_t blocks = (H.T + 63) / 64;
    hipLaunchKernelGGL(k_portfolio_delta, dim3(blocks), dim3(64), 0, s, H);
    return hipGetLastError();
}

}  // namespace nmlmc
