// Synthetic portfolios: portfolio_gen_data / portfolio_free_data
// (portfolio_sampler.cpp:792-912), host only.
//
// The generator is reproduced draw for draw, so that a portfolio generated here
// equals the reference's for the same arguments:
//   - every column is drawn from its own std::mt19937 seeded with `seed` (the
//     reference's std::bind copies the engine per column, so each column starts
//     from the seeded state), with libstdc++'s distributions;
//   - the first Q options are calls and the next Q puts on stocks 0..Q-1;
//   - options are split method-major by floor(probs[m] * P) and priced by
//     portfolio_comp_params; call weights are rescaled per stock to hedge the
//     puts' delta, all weights normalised by their sum, prob = |w| / sqrt(W_est)
//     (or 1), and each method sorted by descending prob with std::sort;
//   - W_est stays 1: the reference's Unbiased work factor is applied to copies
//     of the records (a range-for by value), so it never reaches them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

#include "../../include/nmlmc.h"
#include "host.hpp"

namespace nmlmc {
namespace {

template <class Dist, class T>
void draw_column(uint64_t seed, Dist dist, std::vector<T>& out) {
    std::mt19937 engine(seed);
    for (auto& v : out) v = (T)dist(engine);
}

}  // namespace
}  // namespace nmlmc

using nmlmc::set_error;

extern "C" {

portfolio_data portfolio_gen_data(uint32_t P, uint32_t Q, uint64_t seed, double* probs, bool uniform_weights,
                                  bool uniform_prob) {
    portfolio_data pd;
    std::memset(&pd, 0, sizeof(pd));
    if (!probs || Q == 0 || P < 2 * Q) {
        set_error(NMLMC_ERR_ARG, "portfolio_gen_data needs probs and P >= 2Q > 0");
        return pd;
    }
    uint32_t counts[3], total = 0;
    for (int m = 0; m < 3; m++) {
        counts[m] = (uint32_t)(probs[m] * P);
        total += counts[m];
    }
    if (total != P) {
        set_error(NMLMC_ERR_ARG, "portfolio_gen_data: floor(probs[m] * P) must add up to P");
        return pd;
    }
    pd.r = 0.05;
    pd.rho = 0.2;
    pd.s1_rho2 = std::sqrt(1 - pd.rho * pd.rho);
    pd.max_loss = 0.7;
    pd.tau = 0.02;
    pd.dt_h0inv = 8;
    pd.dt_beta = 4;
    pd.dt_prob_expo = -1.5;
    pd.phi_type = 0;
    pd._phi_size = 1;
    pd.subsample = 0;
    pd.use_delta_cv = true;
    pd.use_coarse_delta_cv = false;
    pd.use_antithetic_risk = true;
    pd.use_milstein = true;

    using real_u = std::uniform_real_distribution<double>;
    using int_u = std::uniform_int_distribution<uint32_t>;
    std::vector<double> S0(Q), sig(Q), mu(Q), maturity(P, 1.), strike(P), weight(P, 0.);
    std::vector<uint32_t> type(P), asset(P);
    nmlmc::draw_column(seed, real_u(90, 110), S0);
    nmlmc::draw_column(seed, real_u(0.01, 0.4), sig);
    nmlmc::draw_column(seed, real_u(0.03, 0.3), mu);
    nmlmc::draw_column(seed, int_u(0, 1), type);
    nmlmc::draw_column(seed, int_u(0, Q - 1), asset);
    nmlmc::draw_column(seed, real_u(1, 5), maturity);
    nmlmc::draw_column(seed, real_u(80, 120), strike);
    if (!uniform_weights) nmlmc::draw_column(seed, std::normal_distribution<double>(0, 1), weight);

    option_params* ops = new option_params[P];
    stock_params* stocks = new stock_params[Q];
    for (uint32_t i = 0; i < Q; i++) {
        type[i] = 0;
        type[Q + i] = 1;
        asset[i] = asset[Q + i] = i;
        stocks[i] = stock_params{S0[i], mu[i], sig[i]};
    }
    for (uint32_t i = 0; i < P; i++) {
        option_params& o = ops[i];
        std::memset(&o, 0, sizeof(o));
        o.w = weight[i];
        o.strike = strike[i];
        o.type = type[i] ? -1 : 1;
        o.maturity = maturity[i];
        o.stock_ind = asset[i];
        o.prob = -1;
        o.W_est = 1;
    }
    pd.stocks.data = stocks;
    pd.stocks.size = Q;
    uint32_t off = 0;
    for (int m = 0; m < 3; m++) {
        pd.options[m].data = ops + off;
        pd.options[m].size = counts[m];
        off += counts[m];
    }
    portfolio_comp_params(&pd);

    // Delta hedge: scale each stock's calls so that their delta offsets the puts'.
    std::vector<double> call_delta(Q, 0.), put_delta(Q, 0.);
    for (uint32_t i = 0; i < P; i++) {
        const double dw = ops[i].delta_est * ops[i].w;
        (ops[i].type == 1 ? call_delta : put_delta)[ops[i].stock_ind] += dw;
    }
    double total_weight = 0;
    for (uint32_t i = 0; i < P; i++) {
        if (ops[i].type == 1) ops[i].w *= -put_delta[ops[i].stock_ind] / call_delta[ops[i].stock_ind];
        total_weight += ops[i].w;
    }
    for (int m = 0; m < 3; m++) {
        option_params* first = (option_params*)pd.options[m].data;
        option_params* last = first + pd.options[m].size;
        for (option_params* o = first; o != last; o++) {
            o->w /= total_weight;
            o->prob = uniform_prob ? 1 : std::abs(o->w) / std::sqrt(o->W_est);
        }
        std::sort(first, last, [](const option_params& x, const option_params& y) { return x.prob > y.prob; });
    }
    set_error(NMLMC_OK, "");
    return pd;
}

void portfolio_free_data(portfolio_data* pdata) {
    if (!pdata) return;
    delete[] (stock_params*)pdata->stocks.data;
    delete[] (option_params*)pdata->options[0].data;
    pdata->stocks.data = nullptr;
    pdata->stocks.size = 0;
    for (int m = 0; m < 3; m++) {
        pdata->options[m].data = nullptr;
        pdata->options[m].size = 0;
    }
}

}  // extern "C"
