// Portfolio sampler device functions (portfolio_sampler.cpp:136-615).
//
// FP64 functions keep the reference's expression order (real_t = double) so the
// plain form agrees with the reference to rounding of the device libm; the
// *_f32 variants are the mixed form's single-precision arithmetic.  Option type
// codes: 1 call, -1 put, 2 / -2 cash-or-nothing digital call / put (paying 1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "portfolio.hpp"
#include "rng.hpp"

namespace nmlmc {

constexpr double kSqrt2 = 1.414213562373095;
constexpr double kSqrt2Pi = 2.506628274631;

__device__ __forceinline__ bool is_digital(int type) { return type == 2 || type == -2; }

// norm_cdf / norm_pdf (common.hpp:201-203).
__device__ __forceinline__ double norm_cdf_d(double x) { return 0.5 * erfc(-x / kSqrt2); }
__device__ __forceinline__ double norm_pdf_d(double x) { return exp(-0.5 * (x * x)) / kSqrt2Pi; }

// simulate_stock (portfolio_sampler.cpp:136-138).
__device__ __forceinline__ double simulate_stock_d(double S0, double mu, double sigma, double t, double dW) {
    return S0 * exp(sigma * dW + (mu - 0.5 * sigma * sigma) * t);
}

// calc_option_blackscholes (140-150); digitals e^{-rt} N(+-d2).
__device__ __forceinline__ double bs_price_d(int type, double S0, double strike, double r, double t, double sigma) {
    const double v1 = log(S0 / strike);
    const double v2 = (r + (sigma * sigma) / 2.0) * t;
    const double v3 = sigma * sqrt(t);
    const double d1 = (v1 + v2) / v3;
    const double d2 = d1 - sigma * sqrt(t);
    if (is_digital(type)) return exp(-r * t) * norm_cdf_d((double)(type / 2) * d2);
    const double mod = (double)type;
    return mod * S0 * norm_cdf_d(mod * d1) - mod * strike * exp(-r * t) * norm_cdf_d(mod * d2);
}

// calc_delta_blackscholes (152-160); digitals +-e^{-rt} n(d2) / (S0 sigma sqrt t).
__device__ __forceinline__ double bs_delta_d(int type, double S0, double strike, double r, double t, double sigma) {
    const double v1 = log(S0 / strike);
    const double v2 = (r + (sigma * sigma) / 2.0) * t;
    const double v3 = sigma * sqrt(t);
    const double d1 = (v1 + v2) / v3;
    if (is_digital(type)) {
        const double d2 = d1 - sigma * sqrt(t);
        return (double)(type / 2) * exp(-r * t) * norm_pdf_d(d2) / (S0 * sigma * sqrt(t));
    }
    return norm_cdf_d(d1) - (double)(type == -1);
}

// Types 3 / -3 (arithmetic-average Asian call / put, an extension; Unbiased
// method only) take the path's average over (tau, T] in place of S_T.
__device__ __forceinline__ bool is_asian(int type) { return type == 3 || type == -3; }
__device__ __forceinline__ double payoff_d(int type, double S, double K) {
    if (type == 2) return (double)(S > K);
    if (type == -2) return (double)(S < K);
    const double v = (double)(is_asian(type) ? type / 3 : type) * (S - K);
    return v > 0 ? v : 0.;
}

// option_diff_from_stock (212-219) and option_derv_from_stock (221-225).
__device__ __forceinline__ double diff_from_stock_d(const OptionRec& o, double S0_T, double St_T, double r) {
    return o.w * exp(-r * o.maturity) * (payoff_d(o.type, S0_T, o.strike) - payoff_d(o.type, St_T, o.strike));
}
__device__ __forceinline__ double derv_from_stock_d(const OptionRec& o, double St_T, double S0, double r) {
    if (is_digital(o.type) || is_asian(o.type)) return 0.;
    const double df = (double)((o.type == 1) - (St_T < o.strike));
    return o.w * (St_T / S0) * exp(-r * o.maturity) * df;
}

// sample_inner_bs (453-468).
template <class RT>
__device__ __forceinline__ double inner_bs_d(const PortfolioDevice& pf, const OptionRec& o, RT R_tau,
                                             double& work) {
    work += o.W_est;
    const uint32_t si = o.stock_ind;
    const StockRec st = pf.stocks[si];
    const double finish = exp(-pf.r * pf.tau) * bs_price_d(o.type, R_tau[si], o.strike, pf.r, o.maturity - pf.tau, st.sigma);
    const double val = o.w * (o.start - finish);
    const double cv = pf.use_delta_cv ? ((R_tau[si] - st.S0) * (o.w * bs_delta_d(o.type, st.S0, o.strike, pf.r, o.maturity, st.sigma))) : 0.;
    return val + cv;
}

// sample_inner_exact (470-519) given its two inner normals.
template <class RT>
__device__ __forceinline__ double inner_exact_d(const PortfolioDevice& pf, const OptionRec& o, RT R_tau,
                                                double z0, double z1, double& work) {
    const bool anti = pf.use_antithetic_risk;
    work += o.W_est;
    const uint32_t si = o.stock_ind;
    const StockRec st = pf.stocks[si];
    const double dW_0_tau = pf.stau * z0;
    const double dW_tau_T = sqrt(o.maturity - pf.tau) * z1;
    const double S0_t_1 = simulate_stock_d(st.S0, pf.r, st.sigma, pf.tau, dW_0_tau);
    const double S0_T_1 = simulate_stock_d(S0_t_1, pf.r, st.sigma, o.maturity - pf.tau, dW_tau_T);
    double S0_T_2 = 0.;
    if (anti) {
        const double S0_t_2 = simulate_stock_d(st.S0, pf.r, st.sigma, pf.tau, -dW_0_tau);
        S0_T_2 = simulate_stock_d(S0_t_2, pf.r, st.sigma, o.maturity - pf.tau, dW_tau_T);
    }
    const double St_T = simulate_stock_d(R_tau[si], pf.r, st.sigma, o.maturity - pf.tau, dW_tau_T);
    const double val1 = diff_from_stock_d(o, S0_T_1, St_T, pf.r);
    const double val2 = anti ? diff_from_stock_d(o, S0_T_2, St_T, pf.r) : val1;
    double dv1 = 0., dv2 = 0.;
    if (pf.use_delta_cv) {
        const double diff = R_tau[si] - st.S0;
        dv1 = diff * derv_from_stock_d(o, S0_T_1, st.S0, pf.r);
        dv2 = anti ? diff * derv_from_stock_d(o, S0_T_2, st.S0, pf.r) : dv1;
    }
    return 0.5 * (val1 + val2 + dv1 + dv2);
}

// simulate_stock_dt (gbm_approx.hpp:18-57) for 3 fine (+ 3 coarse) paths.
// AVG (the Asian extension): A[i] becomes the average of path i's values after
// each of its steps (fine: N steps, coarse: N / coarse_step), in step order.
template <bool AVG>
__device__ __forceinline__ void simulate_dt_d(DStream& rn, double* S, double* A, uint64_t N, double mu,
                                              double sigma, double t, bool with_coarse, bool milstein,
                                              uint32_t coarse_step) {
    const double dt = t / (double)N;
    const double sqrt_dt = sqrt(dt);
    double dW_c = 0.;
    const double dt_c = t / (double)(N / coarse_step);
    // k counts the fine steps since the last coarse step: k reaching
    // coarse_step is (n + 1) % coarse_step == 0, without a 64-bit remainder
    // per step.
    uint32_t k = 0;
    for (uint64_t n = 0; n < N; n++) {
        const double dW = sqrt_dt * rn.randn();
        for (int i = 0; i < 3; i++) {
            const double m = milstein ? (0.5 * (sigma * sigma) * (dW * dW - dt)) : 0.;
            S[i] += S[i] * (mu * dt + sigma * dW + m);
            if (AVG) A[i] += S[i];
        }
        if (with_coarse) {
            dW_c += dW;
            if (++k == coarse_step) {
                k = 0;
                for (int i = 0; i < 3; i++) {
                    const double m = milstein ? (0.5 * (sigma * sigma) * (dW_c * dW_c - dt_c)) : 0.;
                    S[i + 3] += S[i + 3] * (mu * dt_c + sigma * dW_c + m);
                    if (AVG) A[i + 3] += S[i + 3];
                }
                dW_c = 0.;
            }
        }
    }
    if (AVG) {
        for (int i = 0; i < 3; i++) A[i] /= (double)N;
        if (with_coarse)
            for (int i = 0; i < 3; i++) A[i + 3] /= (double)(N / coarse_step);
    }
}

// The Milstein paths of simulate_dt_d split into chunks.  All 3 fine paths
// take the same step S <- S + S g_n (g_n = mu dt + sigma dW_n + Milstein term),
// so a path's value after step n is its starting value times x_n, where
// x_0 = 1, x_n = x_{n-1} + x_{n-1} g_n, and its average over its steps is its
// starting value times (sum of x_n) / N; the coarse paths likewise with the
// sums of beta increments.  A chunk of coarse steps [c0, c1) (fine steps
// [c0 beta, c1 beta)) gives the product of its factors and the sum of its
// partial products, fine and coarse, from unit values; chained in order they
// give the whole paths.  rn stands at the paths' first normal.
struct PathChunk {
    double xf, qf, xc, qc;
};

__device__ __forceinline__ PathChunk path_chunk(DStream rn, uint64_t c0, uint64_t c1, double mu, double sigma,
                                                double dt, double sqrt_dt, double dt_c, uint32_t beta,
                                                bool milstein) {
    rn.skip_normals(c0 * beta);
    PathChunk q{1., 0., 1., 0.};
    for (uint64_t c = c0; c < c1; c++) {
        double dW_c = 0.;
        for (uint32_t i = 0; i < beta; i++) {
            const double dW = sqrt_dt * rn.randn();
            const double m = milstein ? (0.5 * (sigma * sigma) * (dW * dW - dt)) : 0.;
            q.xf += q.xf * (mu * dt + sigma * dW + m);
            q.qf += q.xf;
            dW_c += dW;
        }
        const double m = milstein ? (0.5 * (sigma * sigma) * (dW_c * dW_c - dt_c)) : 0.;
        q.xc += q.xc * (mu * dt_c + sigma * dW_c + m);
        q.qc += q.xc;
    }
    return q;
}

__device__ __forceinline__ uint64_t upow(uint64_t x, uint64_t n) {
    uint64_t r = 1;
    while (n--) r *= x;
    return r;
}

// The randomised level of an Unbiased inner sample (sample_inner_unbiased,
// 521-537): L = floor(ln u / (ln beta * e)) from one sync uniform, so that
// P(L >= l) = beta^(e l).
__device__ __forceinline__ uint32_t unbiased_level(const PortfolioDevice& pf, DStream& sync) {
    return (uint32_t)floor(log(sync.randu()) / pf.unb_lnbe);
}

// One Unbiased inner sample in three parts: unbiased_begin draws its level and
// initial increment and sets the paths' values at tau (3 fine paths: S0 and its
// antithetic twin to tau, and the outer sample's level; the same 3 coarse);
// the paths then run N fine (N / beta coarse) Milstein steps, in one lane
// (simulate_dt_d) or split over a wave; unbiased_end prices the options on them.
struct UnbiasedPath {
    uint32_t L;
    uint64_t N;  // fine steps
    double S[6], A[6];
};

// sample_inner_approx (539-604) up to its call of simulate_stock_dt.
template <class RT>
__device__ __forceinline__ void unbiased_begin(const PortfolioDevice& pf, const OptionRec& o, RT R_tau, DStream& sync,
                                               DStream& inner, double& work, UnbiasedPath& p) {
    p.L = unbiased_level(pf, sync);
    const uint32_t si = o.stock_ind;
    const StockRec st = pf.stocks[si];
    const double dW_0_tau = pf.stau * inner.randn();
    const double S0_t_1 = simulate_stock_d(st.S0, pf.r, st.sigma, pf.tau, dW_0_tau);
    const double S0_t_2 = simulate_stock_d(st.S0, pf.r, st.sigma, pf.tau, -dW_0_tau);
    const uint64_t N_l = upow(pf.dt_beta, p.L);
    p.N = pf.dt_h0inv * N_l;
    work += (double)N_l * o.W_est;
    if (p.L) work += (double)(N_l / pf.dt_beta) * o.W_est;
    const double Rt = R_tau[si];
    p.S[0] = S0_t_1, p.S[1] = Rt, p.S[2] = S0_t_2, p.S[3] = S0_t_1, p.S[4] = Rt, p.S[5] = S0_t_2;
    for (int i = 0; i < 6; i++) p.A[i] = 0.;
}

// The rest of sample_inner_approx after the paths and sample_inner_unbiased's
// weight beta^(-e L) / (1 - beta^e).
template <class RT>
__device__ __forceinline__ double unbiased_end(const PortfolioDevice& pf, const OptionRec& o, RT R_tau,
                                               const UnbiasedPath& p) {
    const double e = pf.dt_prob_expo, x = (double)pf.dt_beta;
    const uint32_t L = p.L;
    const uint32_t si = o.stock_ind;
    const StockRec st = pf.stocks[si];
    // An Asian option's payoffs take the paths' averages over (tau, T].
    const double* V = is_asian(o.type) ? p.A : p.S;
    double val_f = 0., val_c = 0., val_d_f = 0., val_d_c = 0.;
    const double m = pf.use_antithetic_risk ? 0.5 : 1.;
    const double diff = R_tau[si] - st.S0;
    const bool do_d_f = pf.use_delta_cv && (pf.use_coarse_delta_cv || L == 0);
    const bool do_d_c = pf.use_delta_cv && pf.use_coarse_delta_cv;
    for (int i = 0; i < (pf.use_antithetic_risk ? 2 : 1); i++) {
        const int ic = 2 * i;
        val_f += m * diff_from_stock_d(o, V[ic], V[1], pf.r);
        if (do_d_f) val_d_f += m * derv_from_stock_d(o, p.S[ic], st.S0, pf.r);
        if (L == 0) continue;
        val_c += m * diff_from_stock_d(o, V[ic + 3], V[4], pf.r);
        if (do_d_c) val_d_c += m * derv_from_stock_d(o, p.S[ic + 3], st.S0, pf.r);
    }
    const double ret = (val_f - val_c) + diff * (val_d_f - val_d_c);
    const double L_prob_1 = L < kUnbLevels ? pf.unb_weight[L] : pow(x, -e * L) / (1. - pow(x, e));
    return ret * L_prob_1;
}

// sample_inner_unbiased (521-537) + sample_inner_approx (539-604), the paths in
// this lane.
template <class RT>
__device__ __forceinline__ double inner_unbiased_d(const PortfolioDevice& pf, const OptionRec& o, RT R_tau,
                                                   DStream& sync, DStream& inner, double& work) {
    UnbiasedPath p;
    unbiased_begin(pf, o, R_tau, sync, inner, work, p);
    const StockRec st = pf.stocks[o.stock_ind];
    const double T_tau = o.maturity - pf.tau;
    const bool asian = is_asian(o.type);
    if (p.L) {
        if (asian) simulate_dt_d<true>(inner, p.S, p.A, p.N, pf.r, st.sigma, T_tau, true, pf.use_milstein, pf.dt_beta);
        else simulate_dt_d<false>(inner, p.S, p.A, p.N, pf.r, st.sigma, T_tau, true, pf.use_milstein, pf.dt_beta);
    } else {
        if (asian) simulate_dt_d<true>(inner, p.S, p.A, p.N, pf.r, st.sigma, T_tau, false, pf.use_milstein, 2);
        else simulate_dt_d<false>(inner, p.S, p.A, p.N, pf.r, st.sigma, T_tau, false, pf.use_milstein, 2);
    }
    return unbiased_end(pf, o, R_tau, p);
}

// Steps of the level-coupled paths at level l: dt_h0inv 2^floor(l / 2).
__device__ __forceinline__ uint64_t coupled_steps(const PortfolioDevice& pf, uint32_t ell) {
    return (uint64_t)pf.dt_h0inv << (ell / 2);
}

// A level-coupled inner sample of a method-2 option: sample_inner_approx's paths
// (539-604) with a step count set by the level instead of drawn.  The S0 path
// and its antithetic twin go exactly to tau (one normal), then the three paths
// (S0, R_tau, antithetic S0) run n_f Milstein steps over (tau, T]; with `two`,
// also coarse paths of n_f / cs steps on the same increments (cs = 1: the
// coarse paths are the fine ones).  Each value: the antithetic mean of
// w e^{-rT} (payoff(S0 path) - payoff(R_tau path)), on the paths' averages for
// Asians, plus the delta control variate's pathwise term on its own path.
// Returns the fine value, the coarse one in v_c (the fine one without `two`);
// the work is the steps' share of W_est (n / dt_h0inv per path set), as the
// Unbiased method counts it.
template <class RT>
__device__ __forceinline__ double inner_coupled_d(const PortfolioDevice& pf, const OptionRec& o, RT R_tau,
                                                  DStream& inner, uint64_t n_f, uint32_t cs, bool two,
                                                  double& work, double& v_c) {
    const uint32_t si = o.stock_ind;
    const StockRec st = pf.stocks[si];
    const double dW_0_tau = pf.stau * inner.randn();
    const double S0_t_1 = simulate_stock_d(st.S0, pf.r, st.sigma, pf.tau, dW_0_tau);
    const double S0_t_2 = simulate_stock_d(st.S0, pf.r, st.sigma, pf.tau, -dW_0_tau);
    const bool coarse = two && cs > 1;
    work += (double)(n_f / pf.dt_h0inv) * o.W_est;
    if (coarse) work += (double)(n_f / cs / pf.dt_h0inv) * o.W_est;
    const double Rt = R_tau[si];
    double S[6] = {S0_t_1, Rt, S0_t_2, S0_t_1, Rt, S0_t_2};
    double A[6] = {0., 0., 0., 0., 0., 0.};
    const double T_tau = o.maturity - pf.tau;
    const bool asian = is_asian(o.type);
    if (asian) simulate_dt_d<true>(inner, S, A, n_f, pf.r, st.sigma, T_tau, coarse, pf.use_milstein, cs);
    else simulate_dt_d<false>(inner, S, A, n_f, pf.r, st.sigma, T_tau, coarse, pf.use_milstein, cs);
    const double* V = asian ? A : S;
    const double m = pf.use_antithetic_risk ? 0.5 : 1.;
    const double diff = Rt - st.S0;
    double val_f = 0., val_c = 0., val_d_f = 0., val_d_c = 0.;
    for (int i = 0; i < (pf.use_antithetic_risk ? 2 : 1); i++) {
        const int ic = 2 * i;
        val_f += m * diff_from_stock_d(o, V[ic], V[1], pf.r);
        if (pf.use_delta_cv) val_d_f += m * derv_from_stock_d(o, S[ic], st.S0, pf.r);
        if (!coarse) continue;
        val_c += m * diff_from_stock_d(o, V[ic + 3], V[4], pf.r);
        if (pf.use_delta_cv) val_d_c += m * derv_from_stock_d(o, S[ic + 3], st.S0, pf.r);
    }
    const double vf = val_f + diff * val_d_f;
    v_c = coarse ? val_c + diff * val_d_c : vf;
    return vf;
}

// portfolio::sampler::add_phi (portfolio_sampler.cpp:261-291).
__device__ __forceinline__ void phi_add(const PortfolioDevice& pf, double u, double coeff, double* phi) {
    switch (pf.phi_type) {
    case PHI_MAX:
        phi[0] += coeff * (u > 0 ? u : 0.);
        break;
    case PHI_MOMENTS: {
        double p = 1.;
        for (uint32_t i = 0; i < pf.phi_size; i++) {
            p *= u;
            phi[i] += coeff * p;
        }
        break;
    }
    case PHI_FMOMENTS: {
        const double tmp = 3.14159265358979323846 * u;
        for (uint32_t k = 0; k < pf.phi_size / 2; k++) {
            phi[2 * k] += coeff * cos((k + 1) * tmp);
            phi[2 * k + 1] += coeff * sin((k + 1) * tmp);
        }
        if (pf.phi_size % 2) {
            const uint32_t k = pf.phi_size / 2;
            phi[2 * k] += coeff * cos((k + 1) * tmp);
        }
        break;
    }
    default:
        phi[0] += coeff * (double)(u >= 0.);
    }
}

// rand_int_pdf (common.hpp:178-195) over the three compute methods.
__device__ __forceinline__ uint32_t pick_method(const PortfolioDevice& pf, double u) {
    double p = 0.;
    uint32_t ret = 0;
    for (uint32_t i = 0; i < 3; i++) {
        const double c = pf.compute_prob[i];
        if (c <= 0) continue;
        p += c;
        ret = i;
        if (u <= p) break;
    }
    return ret;
}

// Stock levels at tau of one outer sample stored column-wise: stock i of outer
// sample m at p[i * stride] with p = base + m (the batched forms' layout).
struct StridedR {
    const double* p;
    uint64_t stride;
    __device__ __forceinline__ double operator[](uint32_t i) const { return p[(uint64_t)i * stride]; }
};

// rand_int_cdf (common.hpp:160-176).  The reference's fixed-step binary search
// over the exclusive-prefix CDF returns max{i : cdf[i] <= u} (its predicate
// u < cdf[mid] is monotone in mid, and cdf[0] = 0 is never probed).  The same
// index is found here from the guide: the bucket of u bounds the answer by
// [g.x, g.y], which is usually one option (no CDF load at all); otherwise both
// bounds are checked against the CDF (falling back to the whole range if one
// failed) and a binary search closes the gap.
__device__ __forceinline__ uint32_t pick_cdf(const PortfolioDevice& pf, uint32_t m, double u) {
    const uint32_t n = pf.nopt[m];
    if (n == 1) return 0;
    const uint32_t nb = pf.guide_nb[m];
    const double fb = u * pf.guide_scale[m];
    const uint32_t b = fb < (double)(nb - 1) ? (uint32_t)fb : nb - 1;
    const uint32_t* __restrict__ guide = pf.guides[m];
    const uint32_t g = guide[b];
    if (!(g & kGuideMulti)) return g;
    const double* __restrict__ cdf = pf.cdfs[m];
    uint32_t lo = g & kGuideLo, hi = (guide[b + 1] & kGuideLo) + 1;
    if (lo > 0 && !(cdf[lo] <= u)) lo = 0;
    if (hi < n && !(u < cdf[hi])) hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (u < cdf[mid]) hi = mid;
        else lo = mid;
    }
    return lo;
}

// Option choice of sample_inner_random_option_ind (346-370): method from the
// sync role, option from the inner role; returns prob (already over prob_norm).
__device__ __forceinline__ double pick_option(const PortfolioDevice& pf, DStream& sync, DStream& inner,
                                              uint32_t& method, uint32_t& idx) {
    const double uc = sync.randu();
    method = pick_method(pf, uc * pf.prob_norm);
    if (pf.subsample == -1) {
        const double u = inner.randu();
        idx = (uint32_t)(ceil(u * (double)pf.nopt[method]) - 1);
        return 1. / pf.options_count;
    }
    const double u = inner.randu() * pf.compute_prob[method];
    idx = pick_cdf(pf, method, u);
    return pf.options[method][idx].prob / pf.prob_norm;
}

// The draws of pick_option for sub-sampling with probability prob (mode 1),
// without reading the option's probability: method and option index only.
__device__ __forceinline__ uint32_t pick_index(const PortfolioDevice& pf, DStream& st, uint32_t& method) {
    const double uc = st.randu();
    method = pick_method(pf, uc * pf.prob_norm);
    const double u = st.randu() * pf.compute_prob[method];
    return pick_cdf(pf, method, u);
}

// One inner sample of portfolio::sampler::sample_inner (330-334) with the
// reference's draw order.
// UNBIASED = false: for books without Unbiased options, whose code is then left out.
template <bool UNBIASED = true, class RT>
__device__ __forceinline__ double inner_sample_d(const PortfolioDevice& pf, RT R_tau, DStream& sync,
                                                 DStream& inner, double& work) {
    if (pf.subsample == 0) {
        // sample_inner_all_options (336-344).
        double ret = 0.;
        for (uint32_t m = 1; m < (UNBIASED ? 3u : 2u); m++)
            for (uint32_t j = 0; j < pf.nopt[m]; j++) {
                const OptionRec o = pf.options[m][j];
                if (m == 1) {
                    const double z0 = inner.randn();
                    const double z1 = inner.randn();
                    ret += inner_exact_d(pf, o, R_tau, z0, z1, work);
                } else if constexpr (UNBIASED) {
                    ret += inner_unbiased_d(pf, o, R_tau, sync, inner, work);
                }
            }
        return ret;
    }
    uint32_t method, idx;
    const double prob = pick_option(pf, sync, inner, method, idx);
    const OptionRec o = pf.options[method][idx];
    double v;
    if (method == 0) {
        v = inner_bs_d(pf, o, R_tau, work);
    } else if (!UNBIASED || method == 1) {
        const double z0 = inner.randn();
        const double z1 = inner.randn();
        v = inner_exact_d(pf, o, R_tau, z0, z1, work);
    } else {
        v = inner_unbiased_d(pf, o, R_tau, sync, inner, work);
    }
    return v / prob;
}

// init_inner_sampler (312-324): the pre-summed BS options when not sub-sampling.
template <class RT>
__device__ __forceinline__ double det_bs_sum(const PortfolioDevice& pf, RT R_tau, double& work) {
    double det = 0.;
    if (pf.subsample == 0)
        for (uint32_t j = 0; j < pf.nopt[0]; j++) det += inner_bs_d(pf, pf.options[0][j], R_tau, work);
    return det;
}

// ---- FP32 arithmetic of the mixed form (same draws and structure) ----------

__device__ __forceinline__ float simulate_stock_f(float S0, float mu, float sigma, float t, float dW) {
    return S0 * __expf(sigma * dW + (mu - 0.5f * sigma * sigma) * t);
}
__device__ __forceinline__ float payoff_f(int type, float S, float K) {
    if (type == 2) return (float)(S > K);
    if (type == -2) return (float)(S < K);
    const float v = (float)type * (S - K);
    return v > 0.f ? v : 0.f;
}

// sample_inner_exact in single precision; mag receives the sum of the absolute
// values of the terms (the scale of the rounding error).
__device__ __forceinline__ float inner_exact_f(const PortfolioDevice& pf, const OptionRec& o, float R, float z0, float z1,
                                               float& mag, bool& exact64) {
    const bool anti = pf.use_antithetic_risk;
    const StockRec st = pf.stocks[o.stock_ind];
    const float r = (float)pf.r, tau = (float)pf.tau, sig = (float)st.sigma, S0 = (float)st.S0;
    const float K = (float)o.strike, T = (float)o.maturity, w = (float)o.w;
    const float dW0 = (float)pf.stau * z0;
    const float dWT = __fsqrt_rn(T - tau) * z1;
    const float drift_tau = (r - 0.5f * sig * sig) * tau, drift_T = (r - 0.5f * sig * sig) * (T - tau);
    const float gT = __expf(sig * dWT + drift_T);
    const float S1 = S0 * __expf(sig * dW0 + drift_tau) * gT;
    const float S2 = anti ? S0 * __expf(-sig * dW0 + drift_tau) * gT : S1;
    const float St = R * gT;
    const float disc = w * __expf(-r * T);
    const float pt = payoff_f(o.type, St, K);
    const float val1 = disc * (payoff_f(o.type, S1, K) - pt);
    const float val2 = anti ? disc * (payoff_f(o.type, S2, K) - pt) : val1;
    float dv1 = 0.f, dv2 = 0.f;
    const bool cv = pf.use_delta_cv && !is_digital(o.type);
    if (cv) {
        const float diff = R - S0;
        const float c = (float)(o.type == 1);
        const float iS0 = __builtin_amdgcn_rcpf(S0);  // within an ulp or so: inside the term's error scale
        dv1 = diff * disc * (S1 * iS0) * (c - (float)(S1 < K));
        dv2 = anti ? diff * disc * (S2 * iS0) * (c - (float)(S2 < K)) : dv1;
    }
    // Away from the strike every quantity above is a continuous function of
    // the draws and of R: the stock levels carry a relative error of about
    // 2^-20.5 at most (the FP32 normals' error times sigma sqrt(T), plus a few
    // units of 2^-24 from the exponentials of O(1) arguments), so the term's
    // error is at most about 2^-20 times mag.  A digital payoff, or the control
    // variate's indicator S < K, whose stock level lies within 2^-17 K of the
    // strike could jump under that rounding: exact64 then asks for the outer
    // sample to be evaluated in FP64.
    mag = 0.5f * (fabsf(val1) + fabsf(val2) + fabsf(dv1) + fabsf(dv2)) + fabsf(disc) * (S1 + S2 + St);
    const float win = 0x1p-17f * K;
    const bool near1 = fabsf(S1 - K) < win || fabsf(S2 - K) < win;
    exact64 = (is_digital(o.type) && (near1 || fabsf(St - K) < win)) || (cv && near1);
    return 0.5f * (val1 + val2 + dv1 + dv2);
}

// inner_exact_f from an option's MixRec: the term already divided by p and its
// error scale `mag`, such that 2^-18 mag bounds (with a factor 4 to spare) the
// difference between this FP32 term and the FP64 term of the same draws.
// Away from the strike every quantity is a continuous function of the draws and
// of R: the stock levels carry a relative error of about 2^-20.5 at most (the
// FP32 normals' error times sigma sqrt(T), plus a few units of 2^-24 from the
// exponentials of O(1) arguments), so the term's error is at most about 2^-20
// times its continuous scale.  A digital payoff, or the control variate's
// indicator S < K, whose stock level lies within 2^-17 K of the strike may
// jump under that rounding; the jump's size (|discp|, or the indicator's
// factor) enters mag with the weight 2^18, so that the bound covers it whole.
__device__ __forceinline__ float inner_exact_mix(const MixRec& q, const MixStock& st, bool anti, bool cv_on,
                                                 float R, float z0, float z1, float& mag) {
    const int type = (int)(q.stype & 0xffu) - 2;
    // T - tau as the square of the stored root (within an ulp or two of it).
    const float gT = __expf(st.sig * q.sTt * z1 + st.dcoef * (q.sTt * q.sTt));
    const float S1 = st.S0e * __expf(st.sigtau * z0) * gT;
    const float S2 = anti ? st.S0e * __expf(-st.sigtau * z0) * gT : S1;
    const float St = R * gT;
    const float pt = payoff_f(type, St, q.K);
    const float val1 = q.discp * (payoff_f(type, S1, q.K) - pt);
    const float val2 = anti ? q.discp * (payoff_f(type, S2, q.K) - pt) : val1;
    float dv1 = 0.f, dv2 = 0.f;
    const bool digital = is_digital(type);
    const bool cv = cv_on && !digital;
    const float diff = R - st.S0;
    if (cv) {
        const float c = (float)(type == 1);
        dv1 = diff * q.discp * (S1 * st.invS0) * (c - (float)(S1 < q.K));
        dv2 = anti ? diff * q.discp * (S2 * st.invS0) * (c - (float)(S2 < q.K)) : dv1;
    }
    mag = 0.5f * (fabsf(val1) + fabsf(val2) + fabsf(dv1) + fabsf(dv2)) + fabsf(q.discp) * (S1 + S2 + St);
    const float win = 0x1p-17f * q.K;
    const bool near1 = fabsf(S1 - q.K) < win, near2 = fabsf(S2 - q.K) < win, neart = fabsf(St - q.K) < win;
    if (digital) {
        const float at_risk = (float)near1 + (float)near2 + 2.f * (float)neart;
        mag += 0x1p18f * 0.5f * fabsf(q.discp) * at_risk;
    } else if (cv) {
        const float jump = fabsf(diff * q.discp) * (q.K * st.invS0) * 1.001f;
        mag += 0x1p18f * 0.5f * jump * ((float)near1 + (float)near2);
    }
    return 0.5f * (val1 + val2 + dv1 + dv2);
}

}  // namespace nmlmc
