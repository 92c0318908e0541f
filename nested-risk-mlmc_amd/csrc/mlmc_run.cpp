// C entry points of the MLMC driver (mlmc.hpp:290-337) over the GPU level call
// or a caller-supplied level sampler.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nmlmc.h"
#include "host.hpp"
#include "mlmc.hpp"

namespace nmlmc {

namespace {
struct abort_run {
    int code;
};

std::string g_last_table;  // mlmc_output_t::to_string of the last completed run

// Copies a finished run into the C outputs.
void export_run(const mlmc_real::mlmc_output_t& out, double wall, uint32_t max_levels,
                nmlmc_level_stats* levels_out, uint32_t* n_levels, double* summary_out) {
    const uint32_t n = (uint32_t)out.levels.size();
    if (n_levels) *n_levels = n;
    if (levels_out) {
        for (uint32_t l = 0; l < n && l < max_levels; l++) {
            const auto& lv = out.levels[l];
            for (int k = 0; k < 4; k++) {
                levels_out[l].sums_fine[k] = lv.sums_fine[k];
                levels_out[l].sums_diff[k] = lv.sums_diff[k];
            }
            levels_out[l].work = lv.work;
            levels_out[l].time = lv.time;
            levels_out[l].M = lv.M;
        }
    }
    if (summary_out) {
        summary_out[0] = out.starting_level;
        summary_out[1] = out.bias_est;
        summary_out[2] = out.stat_error;
        summary_out[3] = out.TOL;
        summary_out[4] = wall;
    }
}

int run_driver(mlmc_real::mlmc_input_t& in, uint32_t max_levels, nmlmc_level_stats* levels_out,
               uint32_t* n_levels, double* summary_out) {
    const auto t0 = std::chrono::steady_clock::now();
    mlmc_real::mlmc_output_t out;
    try {
        out = mlmc_real::run(in);
    } catch (const abort_run& a) {
        return a.code;
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    g_last_table = out.to_string();
    export_run(out, wall, max_levels, levels_out, n_levels, summary_out);
    return NMLMC_OK;
}

void fill_input(mlmc_real::mlmc_input_t& in, double targetTOL, uint64_t M0, double theta, double Ca,
                double slc, int verbose) {
    in.targetTOL = targetTOL;
    in.M0 = M0;
    in.theta = theta;
    in.Ca = Ca;
    in.starting_level_coeff = slc;
    in.verbose = verbose != 0;
}

// The MLMC drivers of a batch of problems, run together: one host thread per
// instance runs mlmc_t::run unchanged, and its level requests go to the
// calling thread, which waits until every running instance has asked for its
// next step (or finished), then samples each requested level for all of them
// in one batch level call (Engine::batch_sample) and hands the results back.
// Only the calling thread touches the GPU.
struct BatchDriver {
    struct Req {
        std::vector<uint32_t> ells;
        std::vector<uint64_t> Ms;
    };
    std::mutex mu;
    std::condition_variable cv_main, cv_work;
    std::vector<char> waiting, done;
    std::vector<Req> req;
    std::vector<std::vector<mlmc_real::mc_diff_output_t>> res;
    int error = 0;

    std::vector<mlmc_real::mc_diff_output_t> ask(uint32_t i, const std::vector<uint32_t>& ells,
                                                 const std::vector<uint64_t>& Ms) {
        std::unique_lock<std::mutex> lk(mu);
        req[i] = Req{ells, Ms};
        waiting[i] = 1;
        cv_main.notify_one();
        cv_work.wait(lk, [&] { return !waiting[i] || error; });
        if (error) throw abort_run{error};
        return std::move(res[i]);
    }

    // Samples every pending request: one batch level call per requested level.
    int serve(uint32_t n) {
        std::map<uint32_t, std::vector<uint64_t>> levels;
        for (uint32_t i = 0; i < n; i++) {
            if (!waiting[i]) continue;
            res[i].assign(req[i].ells.size(), mlmc_real::mc_diff_output_t());
            for (size_t j = 0; j < req[i].ells.size(); j++) {
                std::vector<uint64_t>& Ms = levels[req[i].ells[j]];
                Ms.resize(n, 0);
                Ms[i] += req[i].Ms[j];
            }
        }
        const size_t w = kMaxPhi * kMaxMoments;
        std::vector<double> fs(n * w), ds(n * w);
        std::vector<sample_output> outs(n);
        for (const auto& lv : levels) {
            std::fill(fs.begin(), fs.end(), 0.);
            std::fill(ds.begin(), ds.end(), 0.);
            for (uint32_t i = 0; i < n; i++) {
                std::memset(&outs[i], 0, sizeof(sample_output));
                outs[i].fsums = &fs[i * w];
                outs[i].diffsums = &ds[i * w];
            }
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = g_engine.batch_sample(lv.first, lv.second.data(), outs.data());
            if (rc) return rc;
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            double wsum = 0.;
            for (uint32_t i = 0; i < n; i++) wsum += outs[i].total_work;
            for (uint32_t i = 0; i < n; i++) {
                if (!waiting[i] || lv.second[i] == 0) continue;
                for (size_t j = 0; j < req[i].ells.size(); j++) {
                    if (req[i].ells[j] != lv.first) continue;
                    mlmc_real::mc_diff_output_t& r = res[i][j];
                    for (int k = 0; k < 4; k++) {
                        r.sums_fine[k] = fs[i * w + k];
                        r.sums_diff[k] = ds[i * w + k];
                    }
                    r.work = outs[i].total_work;
                    r.M = outs[i].M;
                    r.time = wsum > 0. ? dt * outs[i].total_work / wsum : 0.;
                }
            }
        }
        return NMLMC_OK;
    }
};
}  // namespace

}  // namespace nmlmc

using namespace nmlmc;

extern "C" {

const char* nmlmc_mlmc_last_table(void) { return g_last_table.c_str(); }

int nmlmc_mlmc_run(double targetTOL, uint64_t M0, double theta, double Ca, double starting_level_coeff,
                   int verbose, uint32_t max_levels, nmlmc_level_stats* levels_out, uint32_t* n_levels,
                   double* summary_out) {
    if (!g_engine.inited || g_engine.problem < 0)
        return set_error(NMLMC_ERR_STATE, "nmlmc_mlmc_run: no problem staged");
    if (!(targetTOL > 0)) return set_error(NMLMC_ERR_ARG, "targetTOL must be positive");
    mlmc_real::mlmc_input_t in;
    fill_input(in, targetTOL, M0, theta, Ca, starting_level_coeff, verbose);
    const uint32_t ps = g_engine.phi_size();
    // The sample_level lambda of the reference mains (model_sampler.cpp:158-172).
    in.sample_level = [ps](uint32_t ell, uint64_t M) {
        mlmc_real::mc_diff_output_t mc;
        double fs[kMaxPhi * kMaxMoments] = {0}, ds[kMaxPhi * kMaxMoments] = {0};
        sample_output out;
        out.fsums = fs;
        out.diffsums = ds;
        const auto t0 = std::chrono::steady_clock::now();
        int rc = g_engine.sample(ell, M, UINT64_MAX, 0, &out, nullptr, nullptr, nullptr, nullptr);
        if (rc) throw abort_run{rc};
        mc.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        (void)ps;
        for (int k = 0; k < 4; k++) {
            mc.sums_fine[k] = fs[k];
            mc.sums_diff[k] = ds[k];
        }
        mc.work = out.total_work;
        mc.M = out.M;
        return mc;
    };
    // Batched forms: each step's levels as one level sweep (one captured graph,
    // one wait).  The sweep's wall time is shared among its levels in
    // proportion to their work.
    if (g_engine.form != NMLMC_FORM_PLAIN)
        in.sample_levels = [](const std::vector<uint32_t>& ells, const std::vector<uint64_t>& Ms) {
            const size_t n = ells.size(), w = kMaxPhi * kMaxMoments;
            std::vector<double> fs(n * w, 0.), ds(n * w, 0.);
            std::vector<sample_output> outs(n);
            for (size_t i = 0; i < n; i++) {
                std::memset(&outs[i], 0, sizeof(sample_output));
                outs[i].fsums = &fs[i * w];
                outs[i].diffsums = &ds[i * w];
            }
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = g_engine.sample_levels((uint32_t)n, ells.data(), Ms.data(), outs.data());
            if (rc) throw abort_run{rc};
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            double wsum = 0.;
            for (size_t i = 0; i < n; i++) wsum += outs[i].total_work;
            std::vector<mlmc_real::mc_diff_output_t> r(n);
            for (size_t i = 0; i < n; i++) {
                for (int k = 0; k < 4; k++) {
                    r[i].sums_fine[k] = fs[i * w + k];
                    r[i].sums_diff[k] = ds[i * w + k];
                }
                r[i].work = outs[i].total_work;
                r[i].M = outs[i].M;
                r[i].time = wsum > 0. ? dt * outs[i].total_work / wsum : dt / (double)n;
            }
            return r;
        };
    int rc = run_driver(in, max_levels, levels_out, n_levels, summary_out);
    if (rc) return rc;
    return set_error(NMLMC_OK, "");
}

int nmlmc_mlmc_run_cb(nmlmc_sample_level_fn fn, void* user, double targetTOL, uint64_t M0, double theta,
                      double Ca, double starting_level_coeff, int verbose, uint32_t max_levels,
                      nmlmc_level_stats* levels_out, uint32_t* n_levels, double* summary_out) {
    if (!fn) return set_error(NMLMC_ERR_ARG, "sample_level callback is NULL");
    if (!(targetTOL > 0)) return set_error(NMLMC_ERR_ARG, "targetTOL must be positive");
    mlmc_real::mlmc_input_t in;
    fill_input(in, targetTOL, M0, theta, Ca, starting_level_coeff, verbose);
    in.sample_level = [fn, user](uint32_t ell, uint64_t M) {
        nmlmc_level_stats st;
        std::memset(&st, 0, sizeof(st));
        int rc = fn(user, ell, M, &st);
        if (rc) throw abort_run{rc};
        mlmc_real::mc_diff_output_t mc;
        for (int k = 0; k < 4; k++) {
            mc.sums_fine[k] = st.sums_fine[k];
            mc.sums_diff[k] = st.sums_diff[k];
        }
        mc.work = st.work;
        mc.time = st.time;
        mc.M = st.M;
        return mc;
    };
    int rc = run_driver(in, max_levels, levels_out, n_levels, summary_out);
    if (rc) return set_error(NMLMC_ERR_ARG, "sample_level callback returned %d", rc);
    return set_error(NMLMC_OK, "");
}

// A batch of MLMC runs (no counterpart in the reference, which runs one
// problem per process): the instances of the staged batch, instance i to
// targetTOL[i], all with the other driver parameters given; per instance its
// levels at levels_out[i * max_levels ...], the level count at n_levels[i] and
// the summary at summary_out[i * 5 ...] as nmlmc_mlmc_run gives them (wall =
// the whole batch's time).
int nmlmc_batch_mlmc_run(uint32_t n_arg, const double* targetTOL, uint64_t M0, double theta, double Ca,
                         double starting_level_coeff, uint32_t max_levels, nmlmc_level_stats* levels_out,
                         uint32_t* n_levels, double* summary_out) {
    const uint32_t n = (uint32_t)g_engine.batch.size();
    if (n == 0) return set_error(NMLMC_ERR_STATE, "no batch (nmlmc_batch_begin)");
    if (n_arg != n)
        return set_error(NMLMC_ERR_ARG, "batch_mlmc_run: %u array entries for a batch of %u instances", n_arg, n);
    if (!targetTOL) return set_error(NMLMC_ERR_ARG, "targetTOL is NULL");
    for (uint32_t i = 0; i < n; i++)
        if (!(targetTOL[i] > 0)) return set_error(NMLMC_ERR_ARG, "targetTOL[%u] must be positive", i);
    BatchDriver B;
    B.waiting.assign(n, 0);
    B.done.assign(n, 0);
    B.req.resize(n);
    B.res.resize(n);
    std::vector<mlmc_real::mlmc_output_t> outs(n);
    std::vector<int> codes(n, 0);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (uint32_t i = 0; i < n; i++) {
        th.emplace_back([&, i]() {
            mlmc_real::mlmc_input_t in;
            fill_input(in, targetTOL[i], M0, theta, Ca, starting_level_coeff, 0);
            in.sample_levels = [&, i](const std::vector<uint32_t>& ells, const std::vector<uint64_t>& Ms) {
                return B.ask(i, ells, Ms);
            };
            in.sample_level = [&, i](uint32_t ell, uint64_t M) { return B.ask(i, {ell}, {M})[0]; };
            try {
                outs[i] = mlmc_real::run(in);
            } catch (const abort_run& a) {
                codes[i] = a.code;
            }
            std::lock_guard<std::mutex> lk(B.mu);
            B.done[i] = 1;
            B.cv_main.notify_one();
        });
    }
    int rc = NMLMC_OK;
    while (true) {
        std::unique_lock<std::mutex> lk(B.mu);
        B.cv_main.wait(lk, [&] {
            for (uint32_t i = 0; i < n; i++)
                if (!B.done[i] && !B.waiting[i]) return false;
            return true;
        });
        bool any = false;
        for (uint32_t i = 0; i < n; i++) any = any || B.waiting[i];
        if (!any) break;  // every run has finished
        // Every running instance is blocked in ask(): serve them without the lock.
        lk.unlock();
        rc = B.serve(n);
        lk.lock();
        if (rc) B.error = rc;
        for (uint32_t i = 0; i < n; i++) B.waiting[i] = 0;
        B.cv_work.notify_all();
        if (rc) break;
    }
    for (std::thread& t : th) t.join();
    if (rc) return rc;
    for (uint32_t i = 0; i < n; i++)
        if (codes[i]) return codes[i];
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (uint32_t i = 0; i < n; i++)
        export_run(outs[i], wall, max_levels, levels_out ? levels_out + (size_t)i * max_levels : nullptr,
                   n_levels ? n_levels + i : nullptr, summary_out ? summary_out + (size_t)i * 5 : nullptr);
    return set_error(NMLMC_OK, "");
}

}  // extern "C"
