// Probe entry points over the reference's own sources, compiled in place from
// the reference tree by oracle/Makefile into oracle/_ref/.  This file holds no
// reference code: it includes the reference translation units and exposes,
// through a flat C ABI, the pieces a test needs to pin the oracle (RNG blocks,
// stream consumption, sub-sampling searches, Black-Scholes helpers, the MLMC
// driver's call sequence).  TEST INFRASTRUCTURE ONLY: never linked into, or
// called by, the product library.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>
#include <functional>
#include <sstream>
#include <string>

#include "model_sampler.cpp"       // model::sampler, model_* C ABI
#include "portfolio_sampler.cpp"   // portfolio::sampler, portfolio_* C ABI

#ifdef REAL_DOUBLE
#define REFP_REAL_BYTES 8
#else
#define REFP_REAL_BYTES 4
#endif

extern "C" {

int refp_real_bytes() { return REFP_REAL_BYTES; }

// sizeof / offsetof of the structs that cross the reference's C ABI.
int refp_layout(uint64_t* o) {
    int n = 0;
    o[n++] = sizeof(nested_sampler_params);
    o[n++] = offsetof(nested_sampler_params, beta);
    o[n++] = offsetof(nested_sampler_params, antithetic);
    o[n++] = offsetof(nested_sampler_params, fine_moments);
    o[n++] = offsetof(nested_sampler_params, diff_moments);
    o[n++] = offsetof(nested_sampler_params, adaptive_r);
    o[n++] = sizeof(sample_output);
    o[n++] = offsetof(sample_output, fsums);
    o[n++] = offsetof(sample_output, diffsums);
    o[n++] = offsetof(sample_output, total_work);
    o[n++] = offsetof(sample_output, inner_stats);
    o[n++] = sizeof(model::data);
    o[n++] = offsetof(model::data, s1tau);
    o[n++] = offsetof(model::data, max_loss);
    o[n++] = offsetof(model::data, phi_type);
    o[n++] = sizeof(portfolio::option_params);
    o[n++] = offsetof(portfolio::option_params, maturity);
    o[n++] = offsetof(portfolio::option_params, start);
    o[n++] = offsetof(portfolio::option_params, w);
    o[n++] = offsetof(portfolio::option_params, delta_est);
    o[n++] = offsetof(portfolio::option_params, stock_ind);
    o[n++] = offsetof(portfolio::option_params, type);
    o[n++] = offsetof(portfolio::option_params, prob);
    o[n++] = offsetof(portfolio::option_params, W_est);
    o[n++] = sizeof(portfolio::stock_params);
    o[n++] = sizeof(portfolio::data);
    o[n++] = offsetof(portfolio::data, stocks);
    o[n++] = offsetof(portfolio::data, r);
    o[n++] = offsetof(portfolio::data, rho);
    o[n++] = offsetof(portfolio::data, s1_rho2);
    o[n++] = offsetof(portfolio::data, max_loss);
    o[n++] = offsetof(portfolio::data, tau);
    o[n++] = offsetof(portfolio::data, dt_h0inv);
    o[n++] = offsetof(portfolio::data, dt_beta);
    o[n++] = offsetof(portfolio::data, dt_prob_expo);
    o[n++] = offsetof(portfolio::data, phi_type);
    o[n++] = offsetof(portfolio::data, _phi_size);
    o[n++] = offsetof(portfolio::data, subsample);
    o[n++] = offsetof(portfolio::data, use_delta_cv);
    o[n++] = offsetof(portfolio::data, use_coarse_delta_cv);
    o[n++] = offsetof(portfolio::data, use_antithetic_risk);
    o[n++] = offsetof(portfolio::data, use_milstein);
    return n;
}

void refp_set_threads(int n) { omp_set_num_threads(n); }

// Threefry2x64-20 block (Random123/threefry.h:190-293 via r123::Threefry2x64).
void refp_threefry2x64(const uint64_t* ctr, const uint64_t* key, uint64_t* out) {
    r123::Threefry2x64 g;
    r123::Threefry2x64::ctr_type c;
    r123::Threefry2x64::key_type k;
    c[0] = ctr[0]; c[1] = ctr[1];
    k[0] = key[0]; k[1] = key[1];
    auto r = g(c, k);
    out[0] = r[0]; out[1] = r[1];
}

double refp_u01_f64(uint64_t u) { return r123::u01<double>(u); }
double refp_uneg11_f64(uint64_t u) { return r123::uneg11<double>(u); }
float refp_u01_f32(uint64_t u) { return r123::u01<float>(u); }
void refp_boxmuller(uint64_t u0, uint64_t u1, double* out) {
    r123::double2 d = r123::boxmuller(u0, u1);
    out[0] = d.x; out[1] = d.y;
}

// Draw a sequence from one random_state (par_omp.hpp:72-121).
// kinds[i]: 0 = rand() (u64 in out_u), 1 = randu(), 2 = randn() (value in out_r).
void refp_stream(uint64_t seed, uint64_t core, const int* kinds, int n,
                 uint64_t* out_u, double* out_r) {
    random_state s(seed, core);
    for (int i = 0; i < n; i++) {
        out_u[i] = 0; out_r[i] = 0;
        if (kinds[i] == 0) out_u[i] = s.rand();
        else if (kinds[i] == 1) out_r[i] = s.randu();
        else out_r[i] = s.randn();
    }
}

uint32_t refp_rand_int_cdf(double u, const double* cdf, uint32_t n) {
    std::vector<real_t> c(cdf, cdf + n);
    return rand_int_cdf(static_cast<real_t>(u), c.data(), c.data() + n);
}
uint64_t refp_rand_int_pdf(double u, const double* pdf, uint32_t n) {
    std::vector<real_t> c(pdf, pdf + n);
    return rand_int_pdf(static_cast<real_t>(u), c.data(), c.data() + n);
}
uint64_t refp_uint_pow(uint64_t x, uint64_t n) { return uint_pow(x, n); }
double refp_norm_cdf_inv(double p) { return norm_cdf_inv(static_cast<real_t>(p)); }
double refp_norm_cdf(double x) { return norm_cdf(static_cast<real_t>(x)); }

// Black-Scholes helpers (portfolio_sampler.cpp:136-225).
double refp_bs_price(int type, double S0, double K, double r, double t, double sig) {
    return portfolio::calc_option_blackscholes(static_cast<portfolio::OptionType>(type),
                                               S0, K, r, t, sig);
}
double refp_bs_delta(int type, double S0, double K, double r, double t, double sig) {
    return portfolio::calc_delta_blackscholes(static_cast<portfolio::OptionType>(type),
                                              S0, K, r, t, sig);
}
double refp_simulate_stock(double S0, double mu, double sig, double t, double dW) {
    return portfolio::simulate_stock(S0, mu, sig, t, dW);
}

// simulate_stock_dt with a fixed list of normals (gbm_approx.hpp:18-57).
double refp_simulate_stock_dt(int with_coarse, const double* normals, double* S,
                              uint64_t N, double mu, double sig, double t,
                              uint32_t size, int milstein, uint32_t coarse_step) {
    struct fixed_randn {
        const double* p;
        real_t operator()() { return static_cast<real_t>(*p++); }
    };
    std::vector<sum_t> s(S, S + 2 * size);
    real_t W;
    if (with_coarse)
        W = simulate_stock_dt<true>(fixed_randn{normals}, s.data(), N,
                                    static_cast<real_t>(mu), static_cast<real_t>(sig),
                                    static_cast<real_t>(t), size, milstein != 0, coarse_step);
    else
        W = simulate_stock_dt<false>(fixed_randn{normals}, s.data(), N,
                                     static_cast<real_t>(mu), static_cast<real_t>(sig),
                                     static_cast<real_t>(t), size, milstein != 0, coarse_step);
    for (uint32_t i = 0; i < 2 * size; i++) S[i] = s[i];
    return W;
}

// ---- nested sampler driver helpers -------------------------------------------
void refp_init_nested(uint32_t cores, uint32_t N0, uint32_t beta, double adaptive_r,
                      int antithetic, uint32_t fine_moments, uint32_t diff_moments) {
    nested_sampler_params p(fine_moments, diff_moments, N0, beta,
                            static_cast<real_t>(adaptive_r), antithetic != 0);
    if (g_p_nestedmc_data) cleanup_nestedmc_sampler();
    init_nestedmc_sampler(cores, &p);
}

void refp_model_init(int phi_type, double tau, double max_loss, uint64_t seed) {
    model::data d(static_cast<model::PhiType>(phi_type), static_cast<real_t>(tau),
                  static_cast<real_t>(max_loss));
    model::model_init(&d, seed);
}

// One call of model_sample; outputs fsums[moments], diffsums[moments], work, stats[6].
void refp_model_sample(uint32_t ell, uint64_t M, uint32_t moments, double* fsums,
                       double* diffsums, double* work, uint64_t* stats) {
    sample_output out(moments, moments, 1);
    model::model_sample(ell, M, &out);
    for (uint32_t i = 0; i < moments; i++) { fsums[i] = out.fsums[i]; diffsums[i] = out.diffsums[i]; }
    *work = out.total_work;
    for (int i = 0; i < c_inner_stats_count; i++) stats[i] = out.inner_stats[i];
}

}  // extern "C"

// The shift model of the benchmark workloads (outer Y ~ N(0,1), inner
// X = Y + Z, threshold max_loss, Indicator phi) as a sampler class of the
// reference's compile-time SAMPLER concept (the members model::sampler has),
// so that the reference's own engine (sample<>, inner_sampler, random_state;
// nested_sampler.hpp) runs that workload: the CPU baseline of the benchmark.
namespace refp_shift {
static real_t g_max_loss = 2;
class sampler : public inner_sampler_base {
    real_t Y;
public:
    static uint32_t GetWorkspaceSize(const model::data&) { return 0; }
    sampler(void*) : Y(NAN) {}
    void sample_outer(uint64_t, work_t&) { Y = randn_sync(); }
    void copy_outer(const sampler& o) { Y = o.Y; }
    real_t init_inner_sampler(uint64_t&, work_t&) { return -g_max_loss; }
    real_t sample_inner(uint64_t, uint64_t, work_t& w) {
        w++;
        return Y + randn();
    }
    static void add_phi(real_t avg, real_t coeff, real_t* phi) { *phi += coeff * (avg >= 0.f); }
};
}  // namespace refp_shift

extern "C" {

void refp_shift_init(double max_loss, uint64_t seed) {
    refp_shift::g_max_loss = static_cast<real_t>(max_loss);
    gc_phi_size = 1;
    init_samplers<refp_shift::sampler>(seed, 0);
}

void refp_shift_sample(uint32_t ell, uint64_t M, uint32_t moments, double* fsums, double* diffsums,
                       double* work, uint64_t* stats) {
    sample_output out(moments, moments, 1);
    out.M = M;
    sample<refp_shift::sampler>(ell, 0, &out.M, out.fsums, out.diffsums, &out.total_work, out.inner_stats);
    for (uint32_t i = 0; i < moments; i++) { fsums[i] = out.fsums[i]; diffsums[i] = out.diffsums[i]; }
    *work = out.total_work;
    for (int i = 0; i < c_inner_stats_count; i++) stats[i] = out.inner_stats[i];
}

// Portfolio from flat arrays.  Options are given per compute method, already in the
// order the caller wants them staged (descending prob, as portfolio_update_data asserts).
static portfolio::data g_probe_pdata;
static std::vector<portfolio::option_params> g_probe_ops[3];
static std::vector<portfolio::stock_params> g_probe_stocks;

void refp_portfolio_set(uint32_t Q, const double* S0, const double* mu, const double* sigma,
                        const uint32_t* counts,
                        const double* strike, const double* maturity, const double* w,
                        const uint32_t* stock_ind, const int* type, const double* prob,
                        const double* W_est,
                        double r, double rho, double max_loss, double tau,
                        uint32_t dt_h0inv, uint32_t dt_beta, double dt_prob_expo,
                        int phi_type, uint32_t phi_size, int subsample,
                        int use_delta_cv, int use_coarse_delta_cv,
                        int use_antithetic_risk, int use_milstein) {
    g_probe_stocks.resize(Q);
    for (uint32_t i = 0; i < Q; i++) {
        g_probe_stocks[i].S0 = S0[i]; g_probe_stocks[i].mu = mu[i]; g_probe_stocks[i].sigma = sigma[i];
    }
    uint32_t off = 0;
    for (int m = 0; m < 3; m++) {
        g_probe_ops[m].resize(counts[m]);
        for (uint32_t j = 0; j < counts[m]; j++, off++) {
            auto& o = g_probe_ops[m][j];
            o.strike = strike[off]; o.maturity = maturity[off]; o.start = 0; o.w = w[off];
            o.delta_est = 0; o.stock_ind = stock_ind[off];
            o.type = type[off] > 0 ? portfolio::OptionType::Call : portfolio::OptionType::Put;
            o.prob = prob[off]; o.W_est = W_est[off];
        }
        g_probe_pdata.options[m] = portfolio::options_array::create(g_probe_ops[m].data(), counts[m]);
    }
    g_probe_pdata.stocks = portfolio::stocks_array::create(g_probe_stocks.data(), Q);
    g_probe_pdata.r = r; g_probe_pdata.rho = rho;
    g_probe_pdata.s1_rho2 = std::sqrt(1 - POW2(g_probe_pdata.rho));
    g_probe_pdata.max_loss = max_loss; g_probe_pdata.tau = tau;
    g_probe_pdata.dt_h0inv = dt_h0inv; g_probe_pdata.dt_beta = dt_beta;
    g_probe_pdata.dt_prob_expo = dt_prob_expo;
    g_probe_pdata.phi_type = static_cast<portfolio::PhiType>(phi_type);
    g_probe_pdata._phi_size = phi_size;
    g_probe_pdata.subsample = static_cast<portfolio::Subsampling>(subsample);
    g_probe_pdata.use_delta_cv = use_delta_cv; g_probe_pdata.use_coarse_delta_cv = use_coarse_delta_cv;
    g_probe_pdata.use_antithetic_risk = use_antithetic_risk; g_probe_pdata.use_milstein = use_milstein;
    portfolio::portfolio_comp_params(&g_probe_pdata);
}

// Read back start/delta_est computed by portfolio_comp_params.
void refp_portfolio_get_params(double* start, double* delta_est) {
    uint32_t off = 0;
    for (int m = 0; m < 3; m++)
        for (auto& o : g_probe_ops[m]) { start[off] = o.start; delta_est[off] = o.delta_est; off++; }
}

void refp_portfolio_init(uint64_t seed) { portfolio::portfolio_init(&g_probe_pdata, seed); }

void refp_portfolio_sample(uint32_t ell, uint64_t M, uint32_t moments, uint32_t phi_size,
                           double* fsums, double* diffsums, double* work, uint64_t* stats) {
    sample_output out(moments, moments, phi_size);
    portfolio::portfolio_sample(ell, 0, M, &out);
    for (uint32_t i = 0; i < moments * phi_size; i++) { fsums[i] = out.fsums[i]; diffsums[i] = out.diffsums[i]; }
    *work = out.total_work;
    for (int i = 0; i < c_inner_stats_count; i++) stats[i] = out.inner_stats[i];
}

void refp_portfolio_cleanup() { portfolio::portfolio_cleanup(); }

// portfolio_sample_delta_unbiased: the reference copies fine_moments sums
// whatever `moments` is, so it writes into a buffer of the maximum size and
// `moments` values are returned.
void refp_portfolio_sample_delta(uint32_t option_index, uint64_t M, uint32_t moments, double* sums,
                                 double* work) {
    double buf[16] = {0};
    double w = 0;
    portfolio::portfolio_sample_delta_unbiased(option_index, &M, buf, moments, &w);
    for (uint32_t i = 0; i < moments; i++) sums[i] = buf[i];
    *work = w;
}

// portfolio_gen_data, flattened: writes Q stocks and P options in staged order
// (method-major, sorted) plus the scalar fields.
void refp_portfolio_gen(uint32_t P, uint32_t Q, uint64_t seed, const double* probs,
                        int uniform_weights, int uniform_prob,
                        double* S0, double* mu, double* sigma, uint32_t* counts,
                        double* strike, double* maturity, double* start, double* w,
                        double* delta_est, uint32_t* stock_ind, int* type, double* prob,
                        double* W_est) {
    double pr[3] = {probs[0], probs[1], probs[2]};
    portfolio::data d = portfolio::portfolio_gen_data(P, Q, seed, pr, uniform_weights != 0,
                                                      uniform_prob != 0);
    for (uint32_t i = 0; i < Q; i++) { S0[i] = d.stocks[i].S0; mu[i] = d.stocks[i].mu; sigma[i] = d.stocks[i].sigma; }
    uint32_t off = 0;
    for (int m = 0; m < 3; m++) {
        counts[m] = d.options[m].size();
        for (uint32_t j = 0; j < d.options[m].size(); j++, off++) {
            const auto& o = d.options[m][j];
            strike[off] = o.strike; maturity[off] = o.maturity; start[off] = o.start; w[off] = o.w;
            delta_est[off] = o.delta_est; stock_ind[off] = o.stock_ind;
            type[off] = static_cast<int>(o.type); prob[off] = o.prob; W_est[off] = o.W_est;
        }
    }
    portfolio::portfolio_free_data(&d);
}

// ---- MLMC driver call sequence (mlmc.hpp:290-337) with a deterministic level
// sampler: E_l, V_l, W_l given per level; each call returns exact moment sums for
// a two-point distribution with that mean and variance.
struct refp_call { uint32_t ell; uint64_t M; };
static std::vector<refp_call> g_calls;

static std::string g_last_table;

// mlmc_output_t::to_string of the last refp_mlmc_fake_run (mlmc.hpp:106-183).
const char* refp_mlmc_last_table() { return g_last_table.c_str(); }

int refp_mlmc_fake_run(double targetTOL, uint64_t M0, double theta, double Ca,
                       double starting_level_coeff, uint32_t nlev_max,
                       const double* E, const double* V, const double* W,
                       const double* Ef, const double* Vf,
                       uint32_t* out_ell, uint64_t* out_M, int max_calls,
                       double* out_summary) {
    g_calls.clear();
    mlmc_real::mlmc_input_t in;
    in.targetTOL = targetTOL; in.M0 = M0; in.theta = theta; in.Ca = Ca;
    in.starting_level_coeff = starting_level_coeff; in.verbose = false;
    in.sample_level = [&](uint32_t ell, uint64_t M) {
        mlmc_real::mc_diff_output_t o;
        uint32_t l = ell < nlev_max ? ell : nlev_max - 1;
        g_calls.push_back({ell, M});
        // Two-point law: X = mean +- sqrt(var), each with probability 1/2 (M even
        // or odd handled by exact sums of the first M terms of +,-,+,-...).
        auto fill = [&](double mean, double var, real_t* sums) {
            double s = std::sqrt(var);
            uint64_t np = (M + 1) / 2, nm = M / 2;
            double a = mean + s, b = mean - s;
            for (int k = 0; k < 4; k++)
                sums[k] = static_cast<real_t>(np * std::pow(a, k + 1) + nm * std::pow(b, k + 1));
        };
        fill(Ef[l], Vf[l], o.sums_fine);
        fill(E[l], V[l], o.sums_diff);
        o.work = W[l] * M; o.M = M; o.time = 0;
        return o;
    };
    auto out = mlmc_real::run(in);
    g_last_table = out.to_string();
    int n = 0;
    for (auto& c : g_calls) { if (n >= max_calls) break; out_ell[n] = c.ell; out_M[n] = c.M; n++; }
    out_summary[0] = out.levels.size();
    out_summary[1] = out.starting_level;
    out_summary[2] = out.bias_est;
    out_summary[3] = out.stat_error;
    out_summary[4] = out.TOL;
    return static_cast<int>(g_calls.size());
}

}  // extern "C"
