// MLMC driver with the reference's interface and control flow (mlmc.hpp:10-343):
// mlmc_t<T, S> with mc_diff_output_t, mlmc_output_t, mlmc_input_t and run().
// Host-only C++; sample_level is any callable, here the GPU level call.
#pragma once
#include <stdint.h>

#include <cmath>
#include <ctime>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "table.hpp"

namespace nmlmc {

template <typename T>
inline T POW2(T a) { return a * a; }

template <typename T, uint32_t S = 2>
struct mlmc_t {
    static const uint32_t moments = S;

    // Level sufficient statistics (mlmc.hpp:21-85).  Means use the reciprocal
    // (x * (1./M)) and Wl a division, as the reference does.
    struct mc_diff_output_t {
        T sums_fine[S]{};
        T sums_diff[S]{};
        double work{0};
        double time{0};
        uint64_t M{0};

        T dEl() const { return sums_diff[0] * (1. / M); }
        T fEl() const { return sums_fine[0] * (1. / M); }
        double Wl() const { return work / M; }
        T dVl() const { return sums_diff[1] * (1. / M) - POW2(dEl()); }
        T fVl() const { return sums_fine[1] * (1. / M) - POW2(fEl()); }
        // Kurtosis as printed by the reference: its fourth central moment uses the
        // fine sums' second and third moments for both columns (mlmc.hpp:78-84).
        T dKl() const { return mu_4(sums_diff) / POW2(dVl()); }
        T fKl() const { return mu_4(sums_fine) / POW2(fVl()); }

        mc_diff_output_t& operator+=(const mc_diff_output_t& o) {
            for (uint32_t s = 0; s < S; s++) {
                sums_fine[s] += o.sums_fine[s];
                sums_diff[s] += o.sums_diff[s];
            }
            work += o.work;
            time += o.time;
            M += o.M;
            return *this;
        }

    private:
        T mu_4(const T* sums) const {
            T tmp = sums[3] * (1. / M);
            tmp += (sums[0] * (1. / M)) * (sums_fine[2] * (1. / M)) * (-4.);
            tmp += POW2(sums[0] * (1. / M)) * (sums_fine[1] * (1. / M)) * (6);
            tmp += POW2(POW2(sums[0] * (1. / M))) * (-3.);
            return tmp;
        }
    };

    struct mlmc_output_t {
        std::vector<mc_diff_output_t> levels;
        uint32_t starting_level{0};
        double bias_est{std::nan("")};
        double stat_error{std::nan("")};
        double TOL{std::nan("")};

        uint32_t L() const { return levels.size() - starting_level - 1; }

        // total_variance (mlmc.hpp:97-104).
        double total_variance() const {
            uint32_t ell = starting_level;
            double total = levels[ell].fVl() / levels[ell].M;
            for (ell = starting_level + 1; ell < levels.size(); ell++) total += levels[ell].dVl() / levels[ell].M;
            return total;
        }

        // Iteration report in the reference's layout (mlmc.hpp:106-183) so that
        // parse_mlmc_output.py-style readers keep working.
        std::string to_string() const {
            std::stringstream str;
            double discarded_time = 0, discarded_work = 0;
            for (uint32_t ell = 0; ell < starting_level; ell++) {
                discarded_work += levels[ell].work;
                discarded_time += levels[ell].time;
            }
            double total_time = discarded_time, total_work = discarded_work;
            for (uint32_t ell = starting_level; ell < levels.size(); ell++) {
                total_work += levels[ell].work;
                total_time += levels[ell].time;
            }
            T Eg = levels[starting_level].fEl();
            for (uint32_t ell = starting_level + 1; ell < levels.size(); ell++) Eg += levels[ell].dEl();
            str << "TOL: " << TOL << std::endl;
            str << "RMS: " << std::sqrt(POW2(bias_est) + POW2(stat_error)) << " (" << bias_est << ", "
                << stat_error << ") " << std::endl;
            str << "Eg: " << Eg << " --- [" << Eg - TOL << "," << Eg + TOL << "]" << std::endl;
            str << "Discarded Work: " << discarded_work << std::endl;
            str << "Discarded Time: " << discarded_time << std::endl;
            str << "Total Work: " << total_work << std::endl;
            str << "Total Time: " << total_time << std::endl;
            str << std::scientific << std::setprecision(4);
            table_t tab(levels.size());
            tab.add("ell", 3, [&](std::stringstream& o, unsigned i) {
                if (i < starting_level) o << "*" + std::to_string(i);
                else o << i;
            });
            tab.add("dEl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].dEl(); });
            tab.add("fEl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].fEl(); });
            tab.add("dVl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].dVl(); });
            tab.add("fVl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].fVl(); });
            if (S >= 4) {
                tab.add("dKl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].dKl(); });
                tab.add("fKl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].fKl(); });
            }
            tab.add("Wl", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].Wl(); });
            tab.add("M", 5, [&](std::stringstream& o, unsigned i) { o << levels[i].M; });
            tab.add("Time", 10, [&](std::stringstream& o, unsigned i) { o << levels[i].time / levels[i].M; });
            str << tab;
            return str.str();
        }
    };

    // default_bias_estimate (mlmc.hpp:207-214).
    static double default_bias_estimate(const mlmc_output_t& out, double rate = 0.5) {
        if (out.levels.size() > 2) {
            double bias0 = std::abs(out.levels[out.levels.size() - 1].dEl());
            double bias1 = std::abs(out.levels[out.levels.size() - 2].dEl()) * rate;
            return bias0 > bias1 ? bias0 : bias1;
        }
        return std::numeric_limits<double>::infinity();
    }

    // mlmc_input_t (mlmc.hpp:216-239).
    struct mlmc_input_t {
        typedef std::function<mc_diff_output_t(uint32_t ell, uint64_t M)> fn_sample_level_t;
        typedef std::function<double(const mlmc_output_t& out)> fn_est_bias_t;
        mlmc_input_t() { est_bias = [](const mlmc_output_t& out) { return default_bias_estimate(out); }; }
        fn_sample_level_t sample_level;
        // Optional, no counterpart in the reference: the missing samples of
        // every level of one step in one call ((ells, Ms) -> one output per
        // level), e.g. a level sweep captured as one graph.  run() then calls
        // it once per step with the (ell, M) pairs the reference's loop would
        // pass to sample_level one by one, in the same order.
        typedef std::function<std::vector<mc_diff_output_t>(const std::vector<uint32_t>& ells,
                                                            const std::vector<uint64_t>& Ms)>
            fn_sample_levels_t;
        fn_sample_levels_t sample_levels;
        fn_est_bias_t est_bias;
        double targetTOL{1e-3};
        double startTOL{0.5};
        double theta{0.5};
        double starting_level_coeff{1.5};
        double Ca{3};
        bool verbose{true};
        uint64_t M0{100};
        uint32_t zero_protection{0};
    };

    // better_starting_level (mlmc.hpp:241-262).
    static uint32_t better_starting_level(const mlmc_output_t& out, double coeff) {
        uint32_t ell0 = out.starting_level;
        while (true) {
            bool change = false;
            double lhs = std::sqrt(out.levels[ell0].fVl() * out.levels[ell0].Wl());
            for (uint32_t p = ell0 + 1; p < out.levels.size(); p++) {
                lhs += std::sqrt(out.levels[p].dVl() * out.levels[p].Wl());
                double rhs = std::sqrt(out.levels[p].fVl() * out.levels[p].Wl());
                if (lhs / rhs > coeff) {
                    change = true;
                    break;
                }
            }
            if (!change) break;
            ell0 += 1;
        }
        return ell0;
    }

    // get_optimal_Ml (mlmc.hpp:265-288).
    static void get_optimal_Ml(double tol2, uint64_t M0, const mlmc_output_t& out, std::vector<uint64_t>& Ml) {
        uint32_t ell = out.starting_level;
        Ml.resize(out.levels.size());
        std::fill(Ml.begin(), Ml.begin() + ell, 0);
        double lambda = out.levels[ell].M == 0 ? 0 : std::sqrt(out.levels[ell].Wl() * out.levels[ell].fVl());
        for (ell = out.starting_level + 1; ell < out.levels.size(); ell++)
            lambda += out.levels[ell].M == 0 ? 0 : std::sqrt(out.levels[ell].Wl() * out.levels[ell].dVl());
        lambda /= tol2;
        ell = out.starting_level;
        Ml[ell] = (out.levels[ell].M == 0)
                      ? M0
                      : static_cast<uint64_t>(std::ceil(lambda * std::sqrt(out.levels[ell].fVl() / out.levels[ell].Wl())));
        for (ell = out.starting_level + 1; ell < out.levels.size(); ell++)
            Ml[ell] = (out.levels[ell].M == 0)
                          ? M0
                          : static_cast<uint64_t>(std::ceil(lambda * std::sqrt(out.levels[ell].dVl() / out.levels[ell].Wl())));
    }

    // run (mlmc.hpp:290-337): continuation over TOL_k = targetTOL 2^k / sqrt(2)^j.
    static mlmc_output_t run(mlmc_input_t& in) {
        mlmc_output_t out;
        double sqrt_2_inv = 1. / std::sqrt(2);
        uint32_t total_itrs = static_cast<uint32_t>(-(std::log2(in.targetTOL) - std::log2(in.startTOL)));
        out.TOL = in.targetTOL * (1 << total_itrs);
        total_itrs *= 2;
        out.levels.resize(1);
        std::vector<uint64_t> Ml(out.levels.size());
        for (uint32_t itr = 0; itr <= total_itrs; itr++) {
            out.starting_level = better_starting_level(out, in.starting_level_coeff);
            while (true) {
                get_optimal_Ml(in.theta * POW2(out.TOL) / in.Ca, in.M0, out, Ml);
                if (in.sample_levels) {
                    std::vector<uint32_t> ells;
                    std::vector<uint64_t> Ms;
                    for (uint32_t ell = 0; ell < out.levels.size(); ell++)
                        if (out.levels[ell].M < Ml[ell]) {
                            ells.push_back(ell);
                            Ms.push_back(Ml[ell] - out.levels[ell].M);
                            if (in.verbose) std::cout << "# Sampling " << Ms.back() << " of " << ell << std::endl;
                        }
                    if (!ells.empty()) {
                        const std::vector<mc_diff_output_t> r = in.sample_levels(ells, Ms);
                        for (size_t i = 0; i < ells.size(); i++) out.levels[ells[i]] += r[i];
                    }
                }
                for (uint32_t ell = 0; ell < out.levels.size(); ell++) {
                    while (out.levels[ell].M < Ml[ell]) {
                        uint64_t M = Ml[ell] - out.levels[ell].M;
                        if (in.verbose) std::cout << "# Sampling " << M << " of " << ell << std::endl;
                        out.levels[ell] += in.sample_level(ell, M);
                    }
                }
                out.stat_error = std::sqrt(in.Ca * out.total_variance());
                if (POW2(out.stat_error) >= in.theta * POW2(out.TOL)) continue;
                out.bias_est = in.est_bias(out);
                if (POW2(out.bias_est) >= (1 - in.theta) * POW2(out.TOL)) {
                    out.levels.resize(out.levels.size() + 1);
                    continue;
                }
                if (in.verbose) std::cout << out.to_string();
                break;
            }
            out.TOL *= sqrt_2_inv;
        }
        return out;
    }
};

typedef mlmc_t<double, 4> mlmc_real;

}  // namespace nmlmc
