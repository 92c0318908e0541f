// Portfolio entry points (portfolio_sampler.cpp:618-756): host staging of the
// option tables and the level calls of the plain and batched forms.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/nmlmc.h"
#include "host.hpp"

namespace nmlmc {

namespace detail {

// Host restatement of calc_option_blackscholes / calc_delta_blackscholes
// (portfolio_sampler.cpp:140-160) with real_t = double, plus the
// cash-or-nothing digitals of the extension.
const double kSqrt2H = 1.414213562373095;
const double kSqrt2PiH = 2.506628274631;
double norm_cdf_h(double x) { return 0.5 * std::erfc(-x / kSqrt2H); }
double norm_pdf_h(double x) { return std::exp(-0.5 * (x * x)) / kSqrt2PiH; }

double bs_price_h(int type, double S0, double strike, double r, double t, double sigma) {
    const double v1 = std::log(S0 / strike);
    const double v2 = (r + (sigma * sigma) / 2.0) * t;
    const double v3 = sigma * std::sqrt(t);
    const double d1 = (v1 + v2) / v3;
    const double d2 = d1 - sigma * std::sqrt(t);
    if (type == 2 || type == -2) return std::exp(-r * t) * norm_cdf_h((double)(type / 2) * d2);
    const double mod = (double)type;
    return mod * S0 * norm_cdf_h(mod * d1) - mod * strike * std::exp(-r * t) * norm_cdf_h(mod * d2);
}

double bs_delta_h(int type, double S0, double strike, double r, double t, double sigma) {
    const double v1 = std::log(S0 / strike);
    const double v2 = (r + (sigma * sigma) / 2.0) * t;
    const double v3 = sigma * std::sqrt(t);
    const double d1 = (v1 + v2) / v3;
    if (type == 2 || type == -2) {
        const double d2 = d1 - sigma * std::sqrt(t);
        return (double)(type / 2) * std::exp(-r * t) * norm_pdf_h(d2) / (S0 * sigma * std::sqrt(t));
    }
    return norm_cdf_h(d1) - (double)(type == -1);
}

template <class T>
int upload(Engine& e, const T* src, size_t n, const T** dst) {
    void* p = nullptr;
    hipError_t err = hipMalloc(&p, sizeof(T) * (n ? n : 1));
    if (err != hipSuccess) return hip_fail(err, "hipMalloc portfolio table");
    e.pf_allocs.push_back(p);
    if (n) {
        err = hipMemcpy(p, src, sizeof(T) * n, hipMemcpyHostToDevice);
        if (err != hipSuccess) return hip_fail(err, "hipMemcpy portfolio table");
    }
    *dst = (const T*)p;
    return NMLMC_OK;
}

}  // namespace detail

using detail::upload;

void Engine::free_portfolio() {
    for (void* p : pf_allocs) hipFree(p);
    pf_allocs.clear();
    std::memset(&pf, 0, sizeof(pf));
}

// portfolio_update_data (portfolio_sampler.cpp:631-678).
int Engine::stage_portfolio(const portfolio_data* data) {
    if (!inited) return set_error(NMLMC_ERR_STATE, "init_nestedmc_sampler has not been called");
    if (!data) return set_error(NMLMC_ERR_ARG, "portfolio data is NULL");
    tuned.clear();
    const uint32_t d = (uint32_t)data->stocks.size;
    if (d == 0 || !data->stocks.data) return set_error(NMLMC_ERR_ARG, "portfolio has no stocks");
    if (data->subsample < -1 || data->subsample > 1) return set_error(NMLMC_ERR_ARG, "unknown sub-sampling mode");
    if (data->phi_type < 0 || data->phi_type > 3) return set_error(NMLMC_ERR_ARG, "unknown phi type");
    const uint32_t ps = (data->phi_type == 0 || data->phi_type == 1) ? 1u : data->_phi_size;
    if (ps == 0 || ps > (uint32_t)kMaxPhi) return set_error(NMLMC_ERR_ARG, "phi size must be 1..4");
    if (ensure_device()) return nmlmc_last_error();
    free_portfolio();
    PortfolioDevice P;
    std::memset(&P, 0, sizeof(P));
    const stock_params* stocks = (const stock_params*)data->stocks.data;
    std::vector<double> delta(d, 0.);
    P.prob_norm = 0;
    P.total_std = 0;
    P.options_count = 0;
    for (int m = 0; m < 3; m++) {
        const option_params* ops = (const option_params*)data->options[m].data;
        const size_t n = data->options[m].size;
        if (n && !ops) return set_error(NMLMC_ERR_ARG, "options[%d] is NULL", m);
        std::vector<double> cdf(n);
        double c = 0;
        double prev = 1.7976931348623157e308;
        P.compute_prob[m] = 0;
        for (size_t j = 0; j < n; j++) {
            const option_params& o = ops[j];
            if (!(o.prob > 0 && prev >= o.prob))
                return set_error(NMLMC_ERR_ARG, "options[%d][%zu]: probabilities must be positive and "
                                                "sorted in descending order", m, j);
            if (o.stock_ind >= d) return set_error(NMLMC_ERR_ARG, "options[%d][%zu]: stock index out of range", m, j);
            if (o.type != 1 && o.type != -1 && o.type != 2 && o.type != -2 && o.type != 3 && o.type != -3)
                return set_error(NMLMC_ERR_ARG, "options[%d][%zu]: unknown option type", m, j);
            if ((o.type == 3 || o.type == -3) && m != 2)
                return set_error(NMLMC_ERR_ARG, "options[%d][%zu]: Asian options need the Unbiased method (2)",
                                 m, j);
            cdf[j] = c;
            prev = o.prob;
            c += o.prob;
            delta[o.stock_ind] += o.w * o.delta_est;
            P.prob_norm += o.prob;
            P.total_std += o.prob * o.W_est;
            P.compute_prob[m] += o.prob;
        }
        static_assert(sizeof(option_params) == sizeof(OptionRec), "option record layout");
        int rc = upload(*this, (const OptionRec*)ops, n, &P.options[m]);
        if (rc) return rc;
        if ((rc = upload(*this, cdf.data(), n, &P.cdfs[m]))) return rc;
        // Guide of the option search: nb = 4n buckets; bucket b holds the u with
        // (uint32)(u * scale) == b, i.e. [edge(b), edge(b + 1)) with edge(b) the
        // smallest double whose product with scale (one IEEE multiply, as on the
        // device) reaches b; its entry is the answer at both edges.
        if (n) {
            const uint64_t nb = 4 * (uint64_t)n;
            const double scale = (double)nb / c;
            auto edge = [scale](uint64_t b) {
                if (b == 0) return 0.;
                double x = (double)b / scale;
                while (x > 0 && std::nextafter(x, 0.) * scale >= (double)b) x = std::nextafter(x, 0.);
                while (x * scale < (double)b) x = std::nextafter(x, 1e308);
                return x;
            };
            std::vector<uint32_t> at(nb + 1);
            size_t i = 0;
            for (uint64_t b = 0; b < nb; b++) {
                const double x = edge(b);
                while (i + 1 < n && cdf[i + 1] <= x) i++;
                at[b] = (uint32_t)i;
            }
            at[nb] = (uint32_t)(n - 1);  // the last bucket is open above
            if (n > kGuideLo) return set_error(NMLMC_ERR_ARG, "options[%d]: too many options", m);
            std::vector<uint32_t> guide(nb + 1);
            for (uint64_t b = 0; b < nb; b++) guide[b] = at[b] | (at[b + 1] != at[b] ? kGuideMulti : 0u);
            guide[nb] = at[nb];
            if ((rc = upload(*this, guide.data(), nb + 1, &P.guides[m]))) return rc;
            P.guide_scale[m] = scale;
            P.guide_nb[m] = (uint32_t)nb;
        }
        P.nopt[m] = (uint32_t)n;
        P.options_count += (uint32_t)n;
    }
    if (P.options_count == 0) return set_error(NMLMC_ERR_ARG, "portfolio has no options");
    static_assert(sizeof(stock_params) == sizeof(StockRec), "stock record layout");
    int rc = upload(*this, (const StockRec*)stocks, d, &P.stocks);
    if (rc) return rc;
    if ((rc = upload(*this, delta.data(), d, &P.delta_coalesce))) return rc;
    P.chol = nullptr;
    P.d = d;
    P.r = data->r;
    P.rho = data->rho;
    P.s1_rho2 = data->s1_rho2;
    P.max_loss = data->max_loss;
    P.tau = data->tau;
    P.stau = std::sqrt(data->tau);
    P.dt_h0inv = data->dt_h0inv;
    P.dt_beta = data->dt_beta;
    P.dt_prob_expo = data->dt_prob_expo;
    P.unb_lnbe = std::log((double)data->dt_beta) * data->dt_prob_expo;
    for (uint32_t L = 0; L < kUnbLevels; L++)
        P.unb_weight[L] = std::pow((double)data->dt_beta, -data->dt_prob_expo * L) /
                          (1. - std::pow((double)data->dt_beta, data->dt_prob_expo));
    P.phi_type = data->phi_type;
    P.phi_size = ps;
    P.subsample = data->subsample;
    P.use_delta_cv = data->use_delta_cv;
    P.use_coarse_delta_cv = data->use_coarse_delta_cv;
    P.use_antithetic_risk = data->use_antithetic_risk;
    P.use_milstein = data->use_milstein;
    if (data->options[2].size && (data->dt_beta < 2 || !(data->dt_prob_expo < 0) || data->dt_h0inv == 0))
        return set_error(NMLMC_ERR_ARG, "Unbiased options need dt_beta >= 2, dt_prob_expo < 0, dt_h0inv > 0");
    P.coupled = (level_coupled && data->options[2].size) ? 1u : 0u;
    if (P.coupled && data->subsample != 1)
        return set_error(NMLMC_ERR_ARG, "level-coupled paths need sub-sampling mode 1");
    // FP32 records of the Exact options and of the stocks for the mixed form.
    {
        const option_params* ex = (const option_params*)data->options[1].data;
        const size_t n1 = data->options[1].size;
        std::vector<MixRec> mix(n1);
        for (size_t j = 0; j < n1; j++) {
            const option_params& o = ex[j];
            MixRec& q = mix[j];
            q.K = (float)o.strike;
            q.sTt = (float)std::sqrt(o.maturity - data->tau);
            q.discp = (float)(o.w * std::exp(-data->r * o.maturity) * P.prob_norm / o.prob);
            q.stype = (o.stock_ind << 8) | (uint32_t)(o.type + 2);
        }
        if ((rc = upload(*this, mix.data(), n1, &P.mix))) return rc;
        std::vector<MixStock> ms(d);
        for (uint32_t i = 0; i < d; i++) {
            const stock_params& st = stocks[i];
            const double dcoef = data->r - 0.5 * st.sigma * st.sigma;
            MixStock& q = ms[i];
            std::memset(&q, 0, sizeof(q));
            q.sig = (float)st.sigma;
            q.dcoef = (float)dcoef;
            q.sigtau = (float)(st.sigma * std::sqrt(data->tau));
            q.S0e = (float)(st.S0 * std::exp(dcoef * data->tau));
            q.S0 = (float)st.S0;
            q.invS0 = (float)(1. / st.S0);
        }
        if ((rc = upload(*this, ms.data(), d, &P.mix_stocks))) return rc;
    }
    P.w_est_const = -1.;
    {
        bool first = true, same = true;
        double w0 = 0.;
        for (int m = 0; m < 3; m++) {
            const option_params* ops = (const option_params*)data->options[m].data;
            for (size_t j = 0; j < data->options[m].size; j++) {
                if (first) w0 = ops[j].W_est, first = false;
                else same = same && ops[j].W_est == w0;
            }
        }
        if (!first && same) P.w_est_const = w0;
    }
    // Work of one init_inner_sampler call without sub-sampling: the BS options.
    pf_bs_work = 0.;
    if (data->subsample == 0) {
        const option_params* bs = (const option_params*)data->options[0].data;
        for (size_t j = 0; j < data->options[0].size; j++) pf_bs_work += bs[j].W_est;
    }
    pf = P;
    pf_phi_size = ps;
    problem = PROBLEM_PORTFOLIO;
    return set_error(NMLMC_OK, "");
}

int Engine::sample_portfolio_parity(uint32_t ell, uint64_t M, uint32_t, sample_output* out,
                                    double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf, uint64_t* tr_nc) {
    if (pf.coupled)
        return set_error(NMLMC_ERR_STATE, "level-coupled paths run in the batched forms (nmlmc_set_form)");
    PortfolioParityHost H{};
    H.pf = pf;
    H.ep = ep;
    H.cf = level_counts(ep, ell);
    H.cc = level_counts(ep, ell == 0 ? 0 : ell - 1);
    H.seed = seed;
    H.M = M;
    H.T = cores;
    H.ell = ell;
    H.bs_work = pf_bs_work;
    H.blocks = (uint32_t)plain_active(M, cores);
    Partial *parts, *res;
    int rc = plain_buffers(H.blocks, parts, res);
    if (rc) return rc;
    H.states = d_states;
    H.partials = parts;
    TraceBufs tb;
    rc = trace_alloc(M, tr_phi_f || tr_phi_c || tr_nf || tr_nc, tb);
    if (rc) return rc;
    H.tr_phi_f = tb.phi_f;
    H.tr_phi_c = tb.phi_c;
    H.tr_nf = tb.nf;
    H.tr_nc = tb.nc;
    last_G = 64;
    hipError_t e = hipEventRecord(ev0, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = launch_portfolio_parity(H, stream)) != hipSuccess) return hip_fail(e, "launch k_portfolio_parity");
    if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = launch_reduce_partials(parts, H.blocks, res, stream)) != hipSuccess)
        return hip_fail(e, "launch k_reduce_partials");
    last_kernel = "k_portfolio_parity";
    rc = finish_call(ell, out, res);
    if (rc) return rc;
    rc = trace_finish(M, tb, tr_phi_f, tr_phi_c, tr_nf, tr_nc);
    if (rc) return rc;
    return set_error(NMLMC_OK, "");
}

// Stock states at tau of M outer samples (batched forms; grown, never shrunk):
// cap FP64 values, then cap FP32 copies of them for the mixed Exact loop.
int Engine::ensure_pf_R(uint64_t M) { return ensure_R((size_t)M * pf.d); }

int Engine::ensure_R(size_t need) {
    if (need <= pf_R_cap) return NMLMC_OK;
    size_t cap = pf_R_cap ? pf_R_cap : 1024;
    if (pf_R) hipFree(pf_R);
    pf_R = nullptr;
    pf_R32 = nullptr;
    pf_R_cap = 0;
    while (cap < need) cap *= 2;
    const hipError_t e = hipMalloc(&pf_R, (sizeof(double) + sizeof(float)) * cap);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc outer states");
    pf_R32 = reinterpret_cast<float*>(pf_R + cap);
    pf_R_cap = cap;
    return NMLMC_OK;
}

int Engine::sample_portfolio_batched(uint32_t ell, uint64_t M, uint64_t outer_offset, uint32_t lanes_override,
                                     sample_output* out, double* tr_phi_f, double* tr_phi_c, uint64_t* tr_nf,
                                     uint64_t* tr_nc) {
    if (pf.phi_type != PHI_INDICATOR && pf.phi_type != PHI_MAX)
        return set_error(NMLMC_ERR_ARG, "batched forms support the Indicator and Max phi types");
    if (pf.chol && pf.d > 256)
        return set_error(NMLMC_ERR_ARG, "batched forms take a correlation factor of at most 256 stocks");
    PortfolioBatchedHost H{};
    int rc = batched_common(ell, M, outer_offset, H.c);
    if (rc) return rc;
    H.pf = pf;
    H.c.phi_type = pf.phi_type;
    const Shape sh = batched_shape(ell, M, 1, H.c, lanes_override);
    H.c.S = sh.S;
    last_G = H.c.S;
    if ((rc = ensure_pf_R(std::min(M, batched_chunk())))) return rc;
    H.R_tau = pf_R;
    H.bs_work = pf_bs_work;
    H.blocks = (uint32_t)device_cus() * sh.bpc;
    H.mixed = form == NMLMC_FORM_MIXED;
    H.exact_only = pf.subsample == 1 && pf.nopt[0] == 0 && pf.nopt[2] == 0;
    H.multi = false;
    H.R32 = (H.mixed && H.exact_only) ? pf_R32 : nullptr;
    last_kernel = "k_b_final";
    return run_batched(ell, M, H.c, out, [&]() { return launch_portfolio_batched(H, stream); }, tr_phi_f,
                       tr_phi_c, tr_nf, tr_nc);
}

// portfolio_sample_delta_unbiased (portfolio_sampler.cpp:758-781): moments of
// M delta samples of Unbiased option `option_index` from the plain-keyed inner
// streams; reduced in thread order.
int Engine::sample_delta_unbiased(uint32_t option_index, uint64_t M, uint32_t moments, double* sums,
                                  double* work) {
    if (problem != PROBLEM_PORTFOLIO || !d_states) return set_error(NMLMC_ERR_STATE, "no portfolio initialised");
    if (option_index >= pf.nopt[2])
        return set_error(NMLMC_ERR_ARG, "option_index %u: the portfolio has %u Unbiased options", option_index,
                         pf.nopt[2]);
    if (moments == 0 || moments > (uint32_t)kMaxMoments || !sums || !work)
        return set_error(NMLMC_ERR_ARG, "moments must be 1..%d with output buffers", kMaxMoments);
    int rc = ensure_partials(cores);
    if (rc) return rc;
    PortfolioDeltaHost H{};
    H.pf = pf;
    H.seed = seed;
    H.M = M;
    H.T = cores;
    H.option_index = option_index;
    H.moments = moments;
    H.states = d_states;
    H.partials = d_partials;
    hipError_t e = hipEventRecord(ev0, stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = launch_portfolio_delta(H, stream)) != hipSuccess) return hip_fail(e, "launch k_portfolio_delta");
    if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = launch_reduce_partials(d_partials, cores, d_out, stream)) != hipSuccess)
        return hip_fail(e, "launch k_reduce_partials");
    last_kernel = "k_portfolio_delta";
    if ((e = hipMemcpyAsync(h_out, d_out, sizeof(Partial), hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync result");
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "kernel execution");
    float ms = 0;
    hipEventElapsedTime(&ms, ev0, ev1);
    last_ms = ms;
    for (uint32_t i = 0; i < moments; i++) sums[i] = h_out->f[i];
    *work = h_out->work;
    return set_error(NMLMC_OK, "");
}

}  // namespace nmlmc

using nmlmc::g_engine;
using nmlmc::set_error;

extern "C" {

// portfolio_cleanup (portfolio_sampler.cpp:618-629).
void portfolio_cleanup(void) {
    g_engine.free_portfolio();
    if (g_engine.problem == nmlmc::PROBLEM_PORTFOLIO) g_engine.problem = -1;
    set_error(NMLMC_OK, "");
}

// portfolio_update_data (portfolio_sampler.cpp:631-678).
void portfolio_update_data(const portfolio_data* data) { g_engine.stage_portfolio(data); }

// portfolio_init (portfolio_sampler.cpp:680-686).
void portfolio_init(const portfolio_data* data, uint64_t seed) {
    if (g_engine.stage_portfolio(data)) return;
    g_engine.seed_samplers(seed);
}

// portfolio_comp_params (portfolio_sampler.cpp:687-700).  A cash-or-nothing
// digital priced by simulation (Exact / Unbiased) has a zero pathwise delta,
// so its delta estimate is 0: the control variate's constant then matches the
// zero correction its inner samples add.  The same holds for the Asian
// options (types 3 / -3), which have no Black-Scholes value either (start 0:
// the Unbiased method does not use it).
void portfolio_comp_params(portfolio_data* data) {
    if (!data) {
        set_error(NMLMC_ERR_ARG, "portfolio data is NULL");
        return;
    }
    const stock_params* stocks = (const stock_params*)data->stocks.data;
    for (int m = 0; m < 3; m++) {
        option_params* ops = (option_params*)data->options[m].data;
        for (size_t j = 0; j < data->options[m].size; j++) {
            option_params& o = ops[j];
            const stock_params& st = stocks[o.stock_ind];
            if (o.type == 3 || o.type == -3) {
                o.start = 0.;
                o.delta_est = 0.;
                continue;
            }
            o.start = nmlmc::detail::bs_price_h(o.type, st.S0, o.strike, data->r, o.maturity, st.sigma);
            o.delta_est = nmlmc::detail::bs_delta_h(o.type, st.S0, o.strike, data->r, o.maturity, st.sigma);
            if (m > 0 && (o.type == 2 || o.type == -2)) o.delta_est = 0.;
        }
    }
    set_error(NMLMC_OK, "");
}

// portfolio_sample (portfolio_sampler.cpp:749-756).
void portfolio_sample(uint32_t ell, uint64_t override_dyn_inner, uint64_t M, sample_output* out) {
    if (g_engine.problem != nmlmc::PROBLEM_PORTFOLIO) {
        set_error(NMLMC_ERR_STATE, "portfolio_sample: no portfolio staged");
        return;
    }
    g_engine.sample(ell, M, UINT64_MAX, (uint32_t)override_dyn_inner, out, nullptr, nullptr, nullptr, nullptr);
}

// portfolio_sample_delta_unbiased (portfolio_sampler.cpp:758-781).  `sums`
// receives `moments` values (the reference copies fine_moments values into it
// whatever `moments` is); *M is unchanged; *total_work is overwritten.
void portfolio_sample_delta_unbiased(uint32_t option_index, uint64_t* M, double* sums, uint32_t moments,
                                     double* total_work) {
    if (!M) {
        set_error(NMLMC_ERR_ARG, "M is NULL");
        return;
    }
    g_engine.sample_delta_unbiased(option_index, *M, moments, sums, total_work);
}

// portfolio_estimate_delta_for_unbiased (portfolio_sampler.cpp:914-934): for
// each Unbiased option of the staged portfolio, batches of 1000 delta samples
// until 3 sqrt(V / M) < TOL; the mean is added to the caller's delta_est.  The
// loop also stops, with an error status, if the variance is not finite or 2^40
// samples do not reach TOL (the reference would loop for ever).
void portfolio_estimate_delta_for_unbiased(portfolio_data* pdata, double TOL) {
    if (!pdata) {
        set_error(NMLMC_ERR_ARG, "portfolio data is NULL");
        return;
    }
    option_params* ops = (option_params*)pdata->options[2].data;
    for (size_t j = 0; j < pdata->options[2].size; j++) {
        // An Asian option's inner samples add no pathwise delta: its estimate stays 0.
        if (ops[j].type == 3 || ops[j].type == -3) continue;
        uint64_t M = 1000, total_M = 0;
        double total[2] = {0., 0.};
        double sums[2], work = 0.;
        while (true) {
            if (g_engine.sample_delta_unbiased((uint32_t)j, M, 2, sums, &work)) return;
            total[0] += sums[0];
            total[1] += sums[1];
            total_M += M;
            const double mean = total[0] / total_M;
            const double Vl = total[1] / total_M - mean * mean;
            if (3. * std::sqrt(Vl / total_M) < TOL) break;
            if (!std::isfinite(Vl) || total_M >= (1ull << 40)) {
                set_error(NMLMC_ERR_ARG, "delta of Unbiased option %zu does not reach TOL %g", j, TOL);
                return;
            }
        }
        ops[j].delta_est += total[0] / total_M;
    }
    set_error(NMLMC_OK, "");
}

int nmlmc_set_level_coupled_paths(int on) {
    g_engine.level_coupled = on != 0;
    return set_error(NMLMC_OK, "");
}

int nmlmc_portfolio_set_correlation(const double* chol, uint32_t d) {
    if (g_engine.problem != nmlmc::PROBLEM_PORTFOLIO) return set_error(NMLMC_ERR_STATE, "no portfolio staged");
    if (!chol) {
        g_engine.pf.chol = nullptr;
        return set_error(NMLMC_OK, "");
    }
    if (d != g_engine.pf.d) return set_error(NMLMC_ERR_ARG, "correlation factor must be %u x %u", g_engine.pf.d, g_engine.pf.d);
    const double* p = nullptr;
    int rc = nmlmc::detail::upload(g_engine, chol, (size_t)d * d, &p);
    if (rc) return rc;
    g_engine.pf.chol = p;
    return set_error(NMLMC_OK, "");
}

}  // extern "C"
