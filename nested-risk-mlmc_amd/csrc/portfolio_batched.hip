// Batched forms of the portfolio problem: the policy of the batched engine.
//
// Keying as for the model: outer sample g of level l owns the key
// {seed, (l << 46) | g}; its outer draw (portfolio::sampler::sample_outer,
// portfolio_sampler.cpp:293-311, or the Cholesky extension) reads the normal
// stream at counters {j, 0}; inner sample i of draw group `tag` is a
// random_state on counters {j, (tag << 40) | i}, which serves both the sync
// role (method choice, Unbiased level) and the inner role (option choice,
// normals) of that sample in the reference's order.
#include "batched_kernels.hpp"
#include "portfolio.hpp"
#include "portfolio_dev.hpp"

namespace nmlmc {

// EXACT_ONLY: every option is priced by Exact-GBM inner sampling with
// sub-sampling on (the benchmark books); the other methods' code is left out
// of the kernels.  Draws and results are those of the general policy.
// MULTI: a batch of problems in one chain (PortfolioBatchTables): an outer
// sample takes the tables, seed and outer id of its instance, so that its
// draws and results are those of the same outer sample of that problem staged
// alone.
template <bool EXACT_ONLY, bool MULTI = false>
struct PortfolioPolicyT {
    struct Args {
        BatchedCommon c;
        PortfolioDevice pf;
        double* R_tau;  // [M x d] (MULTI: [M x dmax]): stock i of outer sample m at R_tau[m * stride + i]
        double bs_work; // work of one init_inner_sampler call (sub-sampling off)
        PortfolioBatchTables bt;
        float* R32;     // mixed Exact-only form: FP32 copies of R_tau, same layout (else NULL)
    };
    static constexpr uint64_t spu = 1;
    static constexpr bool kMulti = MULTI;
    // The fused first pass only for the Exact-only books: the general policy's
    // inner loop (BS, Exact, Unbiased) is too large to share a kernel with the
    // outer draw and the final at a useful occupancy.
    static constexpr bool kFuse = EXACT_ONLY;
    // The general inner loop (FP64 Unbiased paths beside the FP32 Exact terms)
    // runs without the mixed form's register cap.
    static constexpr bool kMixedWaves = EXACT_ONLY;
    // The general books' level-coupled inner paths (PortfolioDevice::coupled)
    // give the final draws a coarse value too.
    static constexpr bool kTwoValued = !EXACT_ONLY;

    __device__ static const PortfolioDevice& pf_of(const Args& A, const ItemState& it) {
        return MULTI ? A.bt.inst[it.inst] : A.pf;
    }
    __device__ static uint64_t seed_of(const Args& A, const ItemState& it) {
        return MULTI ? A.bt.inst_seed[it.inst] : A.c.seed;
    }
    __device__ static double group_work(const Args& A, const ItemState& it) {
        return MULTI ? A.bt.inst_bs_work[it.inst] : A.bs_work;
    }
    // add_phi (portfolio_sampler.cpp:261-291) for the scalar phi types of the
    // batched forms (Indicator, Max).
    __device__ static double phi(const BatchedCommon& c, double u, double coeff, double acc) {
        return phi_scalar(c.phi_type, u, coeff, acc);
    }
    // MULTI: the segment holding item m (largest k with seg_item0[k] <= m).
    __device__ static uint32_t segment(const Args& A, uint64_t m) {
        uint32_t lo = 0, hi = A.bt.n_seg - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (A.bt.seg_item0[mid] <= m) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }

    // Mixed form: each FP32 inner term differs from its FP64 counterpart by at
    // most 2^-18 of its magnitude (inner_exact_mix; for the other methods' FP32
    // terms, inner_exact_f, a term at a payoff discontinuity makes its segment's
    // magnitude infinite), hence the mean by at most 2^-18 of the mean
    // magnitude.  The BS and Unbiased terms are FP64 in both forms; only the
    // order of their sums differs, which moves a sum of n terms by at most
    // n 2^-53 of their magnitudes: they enter the magnitude at 2^-16 (n up to
    // 2^19).  start is FP64.
    __device__ static double mixed_bound(double mean_mag, double) { return 0x1p-18 * mean_mag; }

    // Mixed form, refinement decisions: bounds on |s1 - s1_64| and |s2 - s2_64|
    // of n terms with error scales summing to sab; smm sums mag_i (|d_i| +
    // 2^-19 mag_i).  An FP32 term d_i is within e_i = 2^-18 mag_i of its FP64
    // value, so |d_i^2 - d64_i^2| <= 2 e_i |d_i| + e_i^2
    // = 2^-17 mag_i (|d_i| + 2^-19 mag_i).  (mag_i, the scale of the stock
    // levels behind the term, is some hundred times |d_i|: the bound on s2 uses
    // |d_i| to stay relative to the variance.)  The FP64 terms' sums differ
    // only in order (covered by sab's 2^-16 share and the caller's 2^-21 s2).
    __device__ static void decision_bounds(double n, double, double sab, double smm, double& e1, double& e2) {
        e1 = 0x1p-18 * sab;
        e2 = 0x1p-17 * smm;
    }

    __device__ static uint64_t key1(const Args& A, uint64_t m) {
        if (MULTI) {
            const uint32_t k = segment(A, m);
            return ((uint64_t)A.c.ell << 46) | (A.bt.seg_outer0[k] + (m - A.bt.seg_item0[k]));
        }
        return ((uint64_t)A.c.ell << 46) | (A.c.offset + m);
    }

    // Stock levels of outer sample m: one row of d values (kept in cache while
    // the outer sample's inner samples gather from it).
    __device__ static uint32_t rstride(const Args& A) { return MULTI ? A.bt.dmax : A.pf.d; }
    __device__ static StridedR stocks_at_tau(const Args& A, uint64_t m) {
        return StridedR{A.R_tau + m * rstride(A), 1};
    }

    // Outer draw (sample_outer, portfolio_sampler.cpp:293-311): the one-factor
    // model here; with a Cholesky factor, k_pf_outer_mfma has already written
    // the stock levels.  Then the delta control variate's constant and, without
    // sub-sampling, the pre-summed BS options (init_inner_sampler, 312-324).
    template <class ACC>
    __device__ static void init_item(const Args& A, uint64_t m, ItemState& it, ACC& acc) {
        it.inst = MULTI ? A.bt.seg_inst[segment(A, m)] : 0;
        const PortfolioDevice& pf = pf_of(A, it);
        double* Rw = A.R_tau + m * rstride(A);
        const uint64_t stride = 1;
        if (pf.chol) {
            acc.st(it.inst, 5, pf.d);
        } else {
            DStream os;
            os.init(seed_of(A, it), key1(A, m), 0, 0, 0., 0);
            const double dW_sys = pf.stau * os.randn();
            float* Rw32 = A.R32 ? A.R32 + m * rstride(A) : nullptr;
            if (pf.d == 1) {
                const StockRec st = pf.stocks[0];
                Rw[0] = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, dW_sys);
                if (Rw32) Rw32[0] = (float)Rw[0];
            } else {
                for (uint32_t i = 0; i < pf.d; i++) {
                    const double dW_ind = pf.stau * os.randn();
                    const double W = pf.rho * dW_sys + pf.s1_rho2 * dW_ind;
                    const StockRec st = pf.stocks[i];
                    const double v = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, W);
                    Rw[(uint64_t)i * stride] = v;
                    if (Rw32) Rw32[(uint64_t)i * stride] = (float)v;
                }
            }
            acc.st(it.inst, 5, os.nn);
        }
        const StridedR R = stocks_at_tau(A, m);
        double cv = 0.;
        // Unrolled so that a lane issues several stocks' loads before it waits
        // (a level with few outer samples has no other work to hide them
        // behind); the additions keep their order.
        if (pf.use_delta_cv) {
#pragma unroll 8
            for (uint32_t i = 0; i < pf.d; i++) cv += (R[i] - pf.stocks[i].S0) * pf.delta_coalesce[i];
        }
        double w0 = 0., det = 0.;
        if (!EXACT_ONLY) det = det_bs_sum(pf, R, w0);  // sub-sampling off: pre-summed BS options
        it.start = det + (-pf.max_loss - cv);
        it.Y = 0.;
    }

    static hipError_t launch_outer(const Args& A, hipStream_t s);

    // The Unbiased samples of level >= 2 of the active lanes' windows (bit j of
    // lvw: unit w0 + j), one at a time, each by all active lanes: its owner
    // lane makes the sample's draws up to its paths (unbiased_begin); the
    // lanes take the chunks of its paths in turn (path_chunk; at most 64 chunks
    // of whole coarse steps, set by the step count alone, so that the value
    // does not depend on how many lanes are active); every lane chains the
    // chunks in order, reading them from the lanes that made them, and the
    // owner prices the option (unbiased_end).
    template <class ADD>
    __device__ static void unbiased_wave(const PortfolioDevice& pf, uint64_t seed, uint64_t k1, uint32_t tag,
                                         uint64_t w0, const StridedR& R, uint64_t lvw, double& work, SegAcc& s,
                                         ADD& add) {
        const uint64_t act = __ballot(1);
        const uint32_t lane = __lane_id();
        const uint32_t rank = (uint32_t)__popcll(act & ((1ull << lane) - 1));
        const uint32_t nact = (uint32_t)__popcll(act);
        for (;;) {
            const uint64_t owners = __ballot(lvw != 0);
            if (owners == 0) break;
            const uint32_t ow = (uint32_t)__builtin_ctzll(owners);
            const bool me = lane == ow;
            DStream st;
            st.init(seed, k1, 0, 0, 0., 0);
            UnbiasedPath p;
            p.N = 0;
            uint32_t idx = 0;
            double prob = 1., sigma = 0., dt = 0., sqrt_dt = 0., dt_c = 0.;
            if (me) {
                const uint32_t j = (uint32_t)__builtin_ctzll(lvw);
                lvw &= lvw - 1;
                st.init(seed, k1, ((uint64_t)tag << 40) | (w0 + j), 0, 0., 0);
                uint32_t method;
                prob = pick_option(pf, st, st, method, idx);
                const OptionRec o = pf.options[2][idx];
                unbiased_begin(pf, o, R, st, st, work, p);
                // simulate_dt_d's step sizes
                const double t = o.maturity - pf.tau;
                sigma = pf.stocks[o.stock_ind].sigma;
                dt = t / (double)p.N;
                sqrt_dt = sqrt(dt);
                dt_c = t / (double)(p.N / pf.dt_beta);
            }
            DStream rn;
            rn.init(bcast_u64(st.k0, ow), bcast_u64(st.k1, ow), bcast_u64(st.c1, ow), bcast_u64(st.P, ow),
                    bcast_f64(st.saved, ow), bcast_u32(st.is_saved, ow));
            const uint64_t N = bcast_u64(p.N, ow);
            const uint32_t beta = bcast_u32(pf.dt_beta, ow);
            const bool milstein = bcast_u32(pf.use_milstein ? 1u : 0u, ow) != 0;
            const double mu = bcast_f64(pf.r, ow), sg = bcast_f64(sigma, ow);
            const double dt_b = bcast_f64(dt, ow), sdt_b = bcast_f64(sqrt_dt, ow), dtc_b = bcast_f64(dt_c, ow);
            const uint64_t K = N / beta;  // coarse steps
            const uint64_t per = (K + 63) / 64;  // coarse steps per chunk
            const uint32_t nch = per ? (uint32_t)((K + per - 1) / per) : 0u;  // <= 64
            double xf = 1., qf = 0., xc = 1., qc = 0.;
            for (uint32_t b0 = 0; b0 < nch; b0 += nact) {
                // Round of chunks [b0, b0 + nact): the lane of rank r takes b0 + r.
                PathChunk q{1., 0., 1., 0.};
                const uint32_t ch = b0 + rank;
                if (ch < nch) {
                    const uint64_t c0 = (uint64_t)ch * per, c1 = (c0 + per < K) ? c0 + per : K;
                    q = path_chunk(rn, c0, c1, mu, sg, dt_b, sdt_b, dtc_b, beta, milstein);
                }
                const uint32_t cnt = (nch - b0 < nact) ? nch - b0 : nact;
                uint64_t lanes = act;  // the lane of rank r is the r-th set bit
                for (uint32_t r = 0; r < cnt; r++) {
                    const uint32_t l = (uint32_t)__builtin_ctzll(lanes);
                    lanes &= lanes - 1;
                    qf += xf * bcast_f64(q.qf, l);
                    xf *= bcast_f64(q.xf, l);
                    qc += xc * bcast_f64(q.qc, l);
                    xc *= bcast_f64(q.xc, l);
                }
            }
            if (me) {
                for (int i = 0; i < 3; i++) {
                    const double S0 = p.S[i];
                    p.S[i] = S0 * xf;
                    p.A[i] = S0 * qf / (double)N;
                    p.S[i + 3] = S0 * xc;
                    p.A[i + 3] = S0 * qc / (double)K;
                }
                add(unbiased_end(pf, pf.options[2][idx], R, p) / prob);
                s.nu += st.nu;
                s.nn += st.nn + N;
            }
        }
    }

    template <bool F32>
    __device__ __forceinline__ static void units(const Args& A, uint32_t m, const ItemState& it, uint32_t tag, uint64_t u0,
                                 uint64_t u1, uint64_t, uint64_t, SegAcc& s) {
        const PortfolioDevice& pf = pf_of(A, it);
        const uint64_t seed = seed_of(A, it);
        const StridedR R = stocks_at_tau(A, m);
        const uint64_t k1 = key1(A, m);
        double a1 = 0., a2 = 0., ab = 0., work = 0.;
        double dc = 0.;  // level-coupled finals: sum of (coarse - fine) values
        Kahan b1, b2, bab;
        float mm = 0.f;
        bool redo64 = false;  // a term at a payoff discontinuity (mixed form)
        if (EXACT_ONLY && F32) {
            // Two inner samples per iteration: both option searches and record
            // loads are issued before either sample's Box-Muller block is
            // computed, so one sample's gathers overlap the other's arithmetic.
            const double cp = pf.compute_prob[1];
            const double wc = pf.w_est_const;
            // The outer sample's stock levels as FP32 (half the bytes of the
            // FP64 row per gather; the same values the FP64 row rounds to).
            const float* R32 = A.R32 ? A.R32 + m * rstride(A) : nullptr;
            auto Rf = [&](uint32_t si) -> float { return R32 ? R32[si] : (float)R[si]; };
            float m2 = 0.f;
            auto term = [&](const MixRec& q, float Rq, const u64x2& blk1, uint32_t idx) {
                float x, y, mag;
                boxmuller_f32(blk1.a, blk1.b, x, y);  // second = word 3 (element a), first = word 2
                const MixStock st = pf.mix_stocks[q.stype >> 8];
                const float d = inner_exact_mix(q, st, pf.use_antithetic_risk, pf.use_delta_cv, Rq, x, y, mag);
                work += wc >= 0. ? wc : pf.options[1][idx].W_est;
                b1.add(d);
                b2.add(d * d);
                bab.add(mag);
                m2 += mag * (fabsf(d) + 0x1p-19f * mag);
            };
            uint64_t u = u0;
            for (; u + 1 < u1; u += 2) {
                const uint64_t ca = ((uint64_t)tag << 40) | u, cb = ca + 1;
                const u64x2 a0 = threefry2x64_20(0, ca, seed, k1);
                const u64x2 b0 = threefry2x64_20(0, cb, seed, k1);
                const uint32_t ia = pick_cdf(pf, 1, u01_f64(a0.a) * cp);
                const uint32_t ib = pick_cdf(pf, 1, u01_f64(b0.a) * cp);
                const MixRec qa = pf.mix[ia], qb = pf.mix[ib];
                const float Ra = Rf(qa.stype >> 8), Rb = Rf(qb.stype >> 8);
                const u64x2 a1 = threefry2x64_20(1, ca, seed, k1);
                const u64x2 b1k = threefry2x64_20(1, cb, seed, k1);
                term(qa, Ra, a1, ia);
                term(qb, Rb, b1k, ib);
            }
            if (u < u1) {
                const uint64_t ca = ((uint64_t)tag << 40) | u;
                const u64x2 a0 = threefry2x64_20(0, ca, seed, k1);
                const uint32_t ia = pick_cdf(pf, 1, u01_f64(a0.a) * cp);
                const MixRec qa = pf.mix[ia];
                term(qa, Rf(qa.stype >> 8), threefry2x64_20(1, ca, seed, k1), ia);
            }
            s.nu += 2 * (u1 - u0);
            s.nn += 2 * (u1 - u0);
            s.s1 += b1.value();
            s.s2 += b2.value();
            s.sab += bab.value();
            s.smm += m2;
            s.work += work;
            return;
        }
        if constexpr (EXACT_ONLY) {
            for (uint64_t u = u0; u < u1; u++) {
                // FP64 (the FP32 loop returned above).  The sample's words: 0 method
                // uniform (only method 1 has options), 1 option uniform (block 0),
                // 2 and 3 the Box-Muller pair (block 1).
                const uint64_t cu = ((uint64_t)tag << 40) | u;
                const u64x2 blk0 = threefry2x64_20(0, cu, seed, k1);
                const u64x2 blk1 = threefry2x64_20(1, cu, seed, k1);
                const uint32_t idx = pick_cdf(pf, 1, u01_f64(blk0.a) * pf.compute_prob[1]);
                s.nu += 2;
                s.nn += 2;
                const OptionRec o = pf.options[1][idx];
                const double prob = o.prob / pf.prob_norm;
                work += o.W_est;
                double z0, z1, w64 = 0.;
                boxmuller_f64(blk1.a, blk1.b, z0, z1);
                const double v = inner_exact_d(pf, o, R, z0, z1, w64) / prob;
                a1 += v;
                a2 += v * v;
            }
        } else if (pf.subsample != 0) {
            // Inner samples by cost.  An Unbiased sample runs 8 * 4^L Milstein
            // steps at its random level L (P(L >= l) = 8^-l), the others a few
            // transcendentals, and a wave whose lanes meet samples of different
            // cost in one iteration waits for the dearest.  So each window of up
            // to 64 units goes first through its BS and Exact samples, noting
            // the level of each Unbiased one; then each lane takes its Unbiased
            // samples of level 0, then those of level 1; then the wave takes the
            // deeper ones one at a time together (unbiased_wave), splitting each
            // sample's paths over its lanes.  A sample's value depends on its own
            // draws only, whichever lanes compute it; only the order of a lane's
            // sums differs from unit order.
            auto add64 = [&](double v) {
                a1 += v;
                a2 += v * v;
                ab += fabs(v);
            };
            // Level-coupled paths (pf.coupled): the draw group's discretisation.
            // Fine tests use the level's step count, coarse tests and the
            // plain estimator's coarse final the level below's; the antithetic
            // final of a level > 0 takes both, on the same increments, and its
            // coarse values go to s1c (dc: their difference from the fine ones).
            const uint32_t ell_g = (tag >= TAG_COARSE_TEST && tag < TAG_FINAL) || tag == TAG_FINAL_COARSE
                                       ? (A.c.ell ? A.c.ell - 1 : 0) : A.c.ell;
            const uint64_t n_cpl = coupled_steps(pf, ell_g);
            const bool two = tag == TAG_FINAL && A.c.ell > 0 && A.c.ep.antithetic;
            const uint32_t cs = two ? (uint32_t)(n_cpl / coupled_steps(pf, A.c.ell - 1)) : 1u;
            for (uint64_t w0 = u0; w0 < u1; w0 += 64) {
                const uint32_t nw = (uint32_t)((u1 - w0) < 64 ? (u1 - w0) : 64);
                uint64_t lv0 = 0, lv1 = 0, lvw = 0;
                for (uint32_t j = 0; j < nw; j++) {
                    DStream st;
                    st.init(seed, k1, ((uint64_t)tag << 40) | (w0 + j), 0, 0., 0);
                    uint32_t method, idx;
                    const double prob = pick_option(pf, st, st, method, idx);
                    if (method == 2 && pf.coupled) {
                        lv0 |= 1ull << j;  // a pass of their own, below
                        continue;
                    }
                    if (method == 2) {
                        const uint32_t L = unbiased_level(pf, st);
                        const uint64_t bit = 1ull << j;
                        if (L == 0) lv0 |= bit;
                        else if (L == 1) lv1 |= bit;
                        else lvw |= bit;
                        continue;
                    }
                    const OptionRec o = pf.options[method][idx];
                    // Mixed form: the Exact terms in FP32, except in books with
                    // Unbiased options, whose FP64 paths dominate the work: there
                    // FP32 Exact terms save little and their error bound sends more
                    // outer samples to the FP64 redo than it saves.
                    if (F32 && method == 1 && pf.nopt[2] == 0) {
                        const uint64_t first = st.rand(), second = st.rand();
                        st.nn += 2;
                        float x, y, mag;
                        boxmuller_f32(second, first, x, y);
                        work += o.W_est;
                        float d;
                        if (pf.subsample == 1) {
                            // The option's FP32 record (its 1/p folded in): a jump
                            // at the strike enters mag instead of asking for FP64.
                            const MixRec q = pf.mix[idx];
                            const uint32_t si = q.stype >> 8;
                            d = inner_exact_mix(q, pf.mix_stocks[si], pf.use_antithetic_risk, pf.use_delta_cv,
                                                (float)R[si], x, y, mag);
                        } else {
                            bool exact64;
                            const float ip = (float)(1. / prob);
                            d = inner_exact_f(pf, o, (float)R[o.stock_ind], x, y, mag, exact64) * ip;
                            mag *= ip;
                            redo64 |= exact64;
                        }
                        b1.add(d);
                        b2.add(d * d);
                        bab.add(mag);
                        mm += mag * (fabsf(d) + 0x1p-19f * mag);
                    } else if (method == 0) {
                        add64(inner_bs_d(pf, o, R, work) / prob);
                    } else {
                        const double z0 = st.randn();
                        const double z1 = st.randn();
                        add64(inner_exact_d(pf, o, R, z0, z1, work) / prob);
                    }
                    s.nu += st.nu;
                    s.nn += st.nn;
                }
                // Unbiased samples of level 0, then level 1, in this lane (or
                // the level-coupled samples, in one pass).
                for (int pass = 0; pass < 2; pass++) {
                    uint64_t mask = pass == 0 ? lv0 : lv1;
                    while (mask) {
                        const uint32_t j = (uint32_t)__builtin_ctzll(mask);
                        mask &= mask - 1;
                        DStream st;
                        st.init(seed, k1, ((uint64_t)tag << 40) | (w0 + j), 0, 0., 0);
                        uint32_t method, idx;
                        const double prob = pick_option(pf, st, st, method, idx);
                        if (pf.coupled) {
                            double vc;
                            const double vf =
                                inner_coupled_d(pf, pf.options[2][idx], R, st, n_cpl, cs, two, work, vc) / prob;
                            vc /= prob;
                            add64(vf);
                            dc += vc - vf;
                            ab += fabs(vc - vf);
                        } else {
                            add64(inner_unbiased_d(pf, pf.options[2][idx], R, st, st, work) / prob);
                        }
                        s.nu += st.nu;
                        s.nn += st.nn;
                    }
                }
                unbiased_wave(pf, seed, k1, tag, w0, R, lvw, work, s, add64);
            }
        } else {
            for (uint64_t u = u0; u < u1; u++) {
                DStream st;
                st.init(seed, k1, ((uint64_t)tag << 40) | u, 0, 0., 0);
                const double d = inner_sample_d(pf, R, st, st, work);
                a1 += d;
                a2 += d * d;
                ab += fabs(d);
                s.nu += st.nu;
                s.nn += st.nn;
            }
        }
        s.s1 += a1 + (F32 ? b1.value() : 0.);
        s.s1c += a1 + dc + (F32 ? b1.value() : 0.);
        s.s2 += a2 + (F32 ? b2.value() : 0.);
        s.sab += F32 ? (redo64 ? (double)INFINITY : bab.value() + 0x1p-16 * ab) : 0.;
        s.smm += F32 ? (double)mm : 0.;
        s.work += work;
    }
};

// Correlated outer draw with a Cholesky factor L: W = sqrt(tau) L z,
// R_i = S0_i exp(sigma_i W_i + (mu_i - sigma_i^2 / 2) tau), as a matrix product
// on the FP64 matrix cores.  A block of 16 waves takes 64 outer samples: their
// d normals (Box-Muller pairs of the outer stream {seed, key}, counter {b, 0}
// for pair b, pairs spread over the waves) go to LDS as z[k][m], rows up to
// the next multiple of 4 zero.  Then wave w forms rows 16 (w % 4) .. +15 of
// Z L^T ([64 x d] x [d x d]) for the 16-stock tiles w / 4, w / 4 + 4, ... with
// v_mfma_f64_16x16x4f64, K in steps of 4 up to the tile's last row (L is lower
// triangular).  FP64 operands and accumulation keep the stock levels those of
// the FP64 forms, which the mixed form's FP64 redo must see too; only the order
// of each row's sum differs from the one-lane restatement's.  Each lane writes
// its accumulator elements: for one element index a wave writes 4 rows x 16
// consecutive stocks.  Sixteen waves per block keep enough loads of L in flight
// when only a few blocks run (the high levels have a few hundred outer samples).
//
// MFMA lane layouts (16x16x4 FP64, checked on gfx950): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; accumulator element i of lane l is D[l / 16 + 4 i][l % 16].
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kOuterWaves = 16;

// A batch of problems (P::kMulti): block b draws 64 outer samples of one
// segment (one instance), whose first block is seg_blk0[k]; blocks of instances
// without a Cholesky factor return at once (their draw is in init_item).
template <class P>
__global__ __launch_bounds__(64 * kOuterWaves) void k_pf_outer_mfma(typename P::Args A) {
    extern __shared__ double zt[];  // [4 ks][64]
    uint64_t m0 = (uint64_t)blockIdx.x * 64, m_end = A.c.M;
    const PortfolioDevice* ppf = &A.pf;
    uint64_t seed = A.c.seed;
    if (P::kMulti) {
        uint32_t lo = 0, hi = A.bt.n_seg - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (A.bt.seg_blk0[mid] <= blockIdx.x) lo = mid;
            else hi = mid - 1;
        }
        m0 = A.bt.seg_item0[lo] + 64 * (uint64_t)(blockIdx.x - A.bt.seg_blk0[lo]);
        m_end = A.bt.seg_item0[lo + 1];
        ppf = &A.bt.inst[A.bt.seg_inst[lo]];
        seed = A.bt.inst_seed[A.bt.seg_inst[lo]];
        if (!ppf->chol) return;
    }
    const PortfolioDevice& pf = *ppf;
    const uint32_t d = pf.d, ks = (d + 3) / 4, nt = (d + 15) / 16, rs = P::rstride(A);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    {
        const uint64_t m = m0 + lane;
        const uint64_t k1 = P::key1(A, m < m_end ? m : m0);
        for (uint32_t b = wv; b < 2 * ks; b += kOuterWaves) {  // pairs 0 .. 2 ks - 1 fill rows 0 .. 4 ks - 1
            double x = 0., y = 0.;
            if (2 * b < d) {
                const u64x2 blk = threefry2x64_20(b, 0, seed, k1);
                boxmuller_f64(blk.a, blk.b, x, y);  // words 2b (element b) and 2b + 1 (element a)
            }
            zt[(2 * b) * 64 + lane] = x;
            zt[(2 * b + 1) * 64 + lane] = 2 * b + 1 < d ? y : 0.;
        }
    }
    __syncthreads();
    const uint32_t mt = wv & 3, ar = 16 * mt + (lane & 15), kq = lane >> 4, col = lane & 15;
    for (uint32_t t = wv >> 2; t < nt; t += kOuterWaves / 4) {
        const uint32_t n = 16 * t + col;
        const uint32_t n_hi = 16 * (t + 1) < d ? 16 * (t + 1) : d;
        const uint32_t s_end = (n_hi + 3) / 4;  // columns k < n_hi
        f64x4 acc = {0., 0., 0., 0.};
        // The elements of L for steps s .. s + 3 are loaded before the products
        // of the previous four wait on them (a level with few outer samples has
        // no other blocks to hide the loads); the products keep their order.
        // Steps past s_end load nothing and multiply nothing (s_end is
        // wave-uniform).
        auto ldb = [&](uint32_t s) -> double {
            const uint32_t k = 4 * s + kq;
            return (s < s_end && n < d && k <= n) ? pf.chol[(size_t)n * d + k] : 0.;
        };
        double b0 = ldb(0), b1 = ldb(1), b2 = ldb(2), b3 = ldb(3);
        for (uint32_t s = 0; s < s_end; s += 4) {
            const double c0 = b0, c1 = b1, c2 = b2, c3 = b3;
            b0 = ldb(s + 4);
            b1 = ldb(s + 5);
            b2 = ldb(s + 6);
            b3 = ldb(s + 7);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(zt[(4 * s + kq) * 64 + ar], c0, acc, 0, 0, 0);
            if (s + 1 < s_end) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(zt[(4 * (s + 1) + kq) * 64 + ar], c1, acc, 0, 0, 0);
            if (s + 2 < s_end) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(zt[(4 * (s + 2) + kq) * 64 + ar], c2, acc, 0, 0, 0);
            if (s + 3 < s_end) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(zt[(4 * (s + 3) + kq) * 64 + ar], c3, acc, 0, 0, 0);
        }
        if (n >= d) continue;
        const StockRec st = pf.stocks[n];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint64_t m = m0 + 16 * mt + (lane >> 4) + 4 * i;
            if (m < m_end) {
                const double v = simulate_stock_d(st.S0, st.mu, st.sigma, pf.tau, pf.stau * acc[i]);
                A.R_tau[m * rs + n] = v;
                if (A.R32) A.R32[m * rs + n] = (float)v;
            }
        }
    }
}

template <bool EXACT_ONLY, bool MULTI>
hipError_t PortfolioPolicyT<EXACT_ONLY, MULTI>::launch_outer(const Args& A, hipStream_t s) {
    uint32_t blocks, dk;
    if (MULTI) {
        blocks = A.bt.outer_blocks;
        dk = A.bt.dmax;
    } else {
        blocks = A.pf.chol ? (uint32_t)((A.c.M + 63) / 64) : 0;
        dk = A.pf.d;
    }
    if (blocks == 0 || A.c.M == 0) return hipSuccess;
    const size_t lds = sizeof(double) * 64 * 4 * ((dk + 3) / 4);  // <= 128 KB for d <= 256
    hipLaunchKernelGGL((k_pf_outer_mfma<PortfolioPolicyT<EXACT_ONLY, MULTI>>), dim3(blocks), dim3(64 * kOuterWaves),
                       lds, s, A);
    return hipGetLastError();
}

template <bool EXACT_ONLY, bool MULTI>
static hipError_t launch_kind(const PortfolioBatchedHost& H, hipStream_t s) {
    using P = PortfolioPolicyT<EXACT_ONLY, MULTI>;
    typename P::Args A{H.c, H.pf, H.R_tau, H.bs_work, H.bt, H.R32};
    return H.mixed ? launch_batched_level<P, true>(A, H.blocks, s) : launch_batched_level<P, false>(A, H.blocks, s);
}

hipError_t launch_portfolio_batched(const PortfolioBatchedHost& H, hipStream_t s) {
    if (H.multi) return H.exact_only ? launch_kind<true, true>(H, s) : launch_kind<false, true>(H, s);
    return H.exact_only ? launch_kind<true, false>(H, s) : launch_kind<false, false>(H, s);
}

}  // namespace nmlmc
