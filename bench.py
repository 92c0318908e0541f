#!/usr/bin/env python3
"""Benchmark of the level-l nested sampler on one MI355X.

Workload (the adaptive model problem): shift model X = Y + Z, Y, Z ~ N(0,1),
threshold 2 (P = 1 - Phi(2)); levels 0..12; base inner count 32 * 2^l, each
outer sample refined by doubling while |mean| < 3 stderr up to 2^6 x; antithetic
fine/coarse estimator; M_l = ceil(1e7 * 2^(-1.5 l)) outer samples per level.

A step is one sweep of sample(l, M_l) over all levels with inputs generated on
the device (counter-based RNG).  value = outer samples / s of the chosen form
(batched FP32+FP64, "mixed") over the timed steps; the plain FP64 form (the
reference keying, 65536 virtual threads) is timed beside it, per level, and
both forms are driven to RMSE eps by the MLMC driver (time to eps).

roofline: a launch is one level call's kernel chain (k_b_zero, k_b_init, one
k_b_test per refinement round, k_b_final_count/alloc/fill, k_b_final,
k_b_redo_split, k_b_redo_seq), which is VALU-bound (integer Threefry + Box-Muller + compensated
sums; no input data, about 1 KB of queue and item-state traffic per outer
sample), so achieved = algorithmic VALU lane-ops per launch (inner samples x ops
per inner sample + outer samples x ops per outer sample, counts in
nested-risk-mlmc_amd/opcount.json) / the chain's duration from HIP events
recorded on the library's stream around it; peak = 256 CU x 64 lanes x 2.4 GHz
= 39.3 T lane-ops/s.

traffic: HBM bytes per launch (level call) from two rocprofv3 --pmc
passes (FETCH_SIZE, WRITE_SIZE), each scaled by what the same counter reads for
a plain device copy of known size (python bench.py --copy-calibration); the
scaled figures are kept in profiles/hbm_traffic.json and read from there.

cpu_baseline: the oracle (C restatement, FP64, same keying and workload) on the
host cores with OpenMP, over a bounded sample of the sweep (M_l scaled down).
Its calibration against the reference's own OpenMP engine (oracle/_ref, where
present) on the workload both run (the non-adaptive model sweep below) is
reported beside it.

Also reported: per-level comparisons of the mixed form with the batched FP64
form on the same outer samples (count flips, indicator flips, FP64 redos and
rechecked decisions, level-mean differences), for the model and for the large
book; the portfolio legs' rooflines (VALU lane-ops of the Exact inner samples,
opcount.json); the non-adaptive model sweep (the reference's CPU path, levels
0..8, M_l = ceil(1e6 2^-1.5l)) on the GPU forms and on the reference's own
OpenMP engine; the small book beside the large one.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_VALU_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.3 T lane-ops/s


def level_M(M0, ell):
    return int(math.ceil(M0 * 2.0 ** (-1.5 * ell)))


def setup(nm, form, cores, seed=1):
    nm.init_nestedmc_sampler(cores, 32, 2, 0.0, True)
    nm.set_adaptive_rule(1, 6)
    nm.set_form(form)
    nm.model_init_ex(nm.MODEL_SHIFT, nm.PHI_INDICATOR, 0.5, 2.0, seed)


def sweep(nm, M0, levels, rec=None):
    for ell in range(levels + 1):
        M = level_M(M0, ell)
        t0 = time.perf_counter()
        o = nm.sample(ell, M)
        dt = time.perf_counter() - t0
        if rec is not None:
            reev, flips, rck = nm.mixed_stats()
            rec.append(dict(ell=ell, M=M, wall=dt, kernel_ms=nm.last_kernel_ms(),
                            inner=int(o["stats"][0]), G=nm.last_group_lanes(), reev=reev,
                            flips=flips, rechecks=rck, fmean=o["fsums"][0] / M,
                            dmean=(o["diffsums"][0] / M) if ell else 0.))


def compare_forms(nm, setup_fn, M0, levels):
    """The mixed and the batched FP64 forms on the same outer samples (outer ids
    0..M_l - 1 of every level): per level the refinement-count flips (N_f or
    N_c differing), indicator flips (phi differing), the mixed form's FP64
    redos, flips of its FP32 indicators before the redo and decisions
    rechecked in FP64, and the difference of the level means (fine mean at
    l = 0, difference mean above)."""
    runs = {}
    for form in (nm.FORM_MIXED, nm.FORM_BATCHED):
        setup_fn(nm, form, 1)
        rs = []
        for ell in range(levels + 1):
            M = level_M(M0, ell)
            o = nm.sample(ell, M, outer_offset=0, trace=True)
            o["mixed_stats"] = nm.mixed_stats() if form == nm.FORM_MIXED else (0, 0, 0)
            rs.append(o)
        nm.cleanup_nestedmc_sampler()
        runs[form] = rs
    out = []
    for ell, (a, b) in enumerate(zip(runs[nm.FORM_MIXED], runs[nm.FORM_BATCHED])):
        M = level_M(M0, ell)
        ma = (a["fsums"][0] if ell == 0 else a["diffsums"][0]) / M
        mb = (b["fsums"][0] if ell == 0 else b["diffsums"][0]) / M
        reev, flips, rck = a["mixed_stats"]
        out.append({"ell": ell, "M": M,
                    "count_flips": int(np.sum((a["nf"] != b["nf"]) | (a["nc"] != b["nc"]))),
                    "indicator_flips": int(np.sum((a["phi_f"] != b["phi_f"]) | (a["phi_c"] != b["phi_c"]))),
                    "redone_fp64": int(reev), "fp32_indicator_flips_before_redo": int(flips),
                    "rechecked_decisions": int(rck), "level_mean_mixed": ma, "level_mean_fp64": mb,
                    "level_mean_diff": abs(ma - mb)})
    return out


def autotune(nm, M0, levels):
    """Launch-shape autotuning of every level call of the sweep, before the
    timed region (the tuning calls draw outer samples set apart for tuning):
    [segment, blocks per CU] per level."""
    return [list(nm.autotune_level(l, level_M(M0, l))[:2]) for l in range(levels + 1)]


def ops_per_sample(form):
    with open(os.path.join(ROOT, "nested-risk-mlmc_amd", "opcount.json")) as f:
        oc = json.load(f)
    key = "mixed" if form == 2 else "fp64"
    return oc["shift_model"][key]["inner"], oc["shift_model"][key]["outer"]


def traffic_bytes(kernel):
    """HBM bytes per launch of ``kernel`` measured by rocprofv3 --pmc and scaled
    by the copy calibration, or None when no measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return rec["bytes_per_launch"]


def copy_calibration(nbytes=1 << 30, reps=4):
    """A plain device-to-device copy of known size for scaling FETCH_SIZE and
    WRITE_SIZE (torch is only device-memory plumbing here)."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "kernel": "copy of %d float32" % n}))


def roofline(recs, form):
    oi, oo = ops_per_sample(form)
    ops = sum(r["inner"] * oi + r["M"] * oo for r in recs)
    secs = sum(r["kernel_ms"] for r in recs) / 1e3
    achieved = ops / secs / 1e12
    chain = "batched_level_chain_mixed" if form == 2 else None
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_VALU_TOPS, 1),
            "unit": "TOP/s", "frac": round(achieved / PEAK_VALU_TOPS, 4),
            "traffic": traffic_bytes(chain) if chain else None,
            "kernel": ("level-call kernel chain k_b_zero..k_b_redo_seq (mixed)" if form == 2 else
                       "batched level-call kernel chain" if form else "k_model_parity"),
            "launches": len(recs), "avg_launch_ms": round(1e3 * secs / len(recs), 4),
            "ops_per_inner_sample": oi}


def portfolio_roofline(recs, form_key):
    """Roofline of the large book's level calls: Exact inner samples x VALU
    lane-ops per sample (opcount.json, portfolio_exact) over the level calls'
    kernel time (the outer draws and the engine's bookkeeping are not counted
    as algorithmic work)."""
    with open(os.path.join(ROOT, "nested-risk-mlmc_amd", "opcount.json")) as f:
        ops = json.load(f)["portfolio_exact"][form_key]["inner"]
    inner = sum(r["inner"] for r in recs)
    secs = sum(r["kernel_ms"] for r in recs) / 1e3
    achieved = inner * ops / secs / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_VALU_TOPS, 1), "unit": "TOP/s",
            "frac": round(achieved / PEAK_VALU_TOPS, 4), "ops_per_inner_sample": ops,
            "inner_samples_per_s": round(inner / secs, 1), "launches": len(recs),
            "avg_launch_ms": round(1e3 * secs / len(recs), 4)}


def per_level(recs, levels):
    out = []
    for ell in range(levels + 1):
        rs = sorted(r["wall"] for r in recs if r["ell"] == ell)
        M = next(r["M"] for r in recs if r["ell"] == ell)
        out.append(round(M / rs[len(rs) // 2], 1))
    return out


def cpu_baseline(M0_cpu, levels):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle  # the checker, timed as the CPU baseline
    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    orc = Oracle("f64").new(1, 32, 2, 0.0, True, adapt_rule=1, cap_log2=6)
    orc.model_init(1, 0, 0.5, 2.0, 1)
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(threads)
    except OSError:
        pass
    total = 0
    t0 = time.perf_counter()
    for ell in range(levels + 1):
        M = level_M(M0_cpu, ell)
        orc.sample(ell, M, keying=1, outer_offset=0)
        total += M
    dt = time.perf_counter() - t0
    return {"value": round(total / dt, 1), "unit": "outer samples/s", "cores": threads,
            "kind": "port",
            "sample": "one level sweep l=0..%d with M_l = ceil(%g * 2^-1.5l) (%d outer samples, "
                      "%.1f s), FP64, batched keying, OpenMP over fixed 256-sample chunks"
                      % (levels, M0_cpu, total, dt)}


def cpu_nonadaptive_sweep(M0, levels=8):
    """Config 1 on the host cores: the non-adaptive shift-model sweep (N = 32 2^l,
    antithetic, seed 1, M_l = ceil(M0 2^-1.5l)) with the reference keying, on
    the reference's own OpenMP engine (oracle/_ref, where it was built) and on
    the oracle, with the same threads: outer samples/s of each and the ratio
    (the calibration of the oracle as the CPU baseline)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes
    from oracle import Oracle, Ref, ref_available  # the checker, timed as the CPU baseline
    threads = min(16, os.cpu_count() or 1)
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(threads)
    except OSError:
        pass

    def sweep(fn):
        per, total = [], 0
        t0 = time.perf_counter()
        for ell in range(levels + 1):
            M = level_M(M0, ell)
            t1 = time.perf_counter()
            fn(ell, M)
            per.append(round(M / (time.perf_counter() - t1), 1))
            total += M
        return round(total / (time.perf_counter() - t0), 1), per

    def best(runs):
        # The best rate over the runs, overall and per level.
        return (max(r for r, _ in runs),
                [max(v) for v in zip(*(p for _, p in runs))])

    # Each sweep continues the streams of the one before (the same work).  The
    # two engines' sweeps alternate, so that a change of the host's load during
    # the measurement falls on both.
    orc = Oracle("f64").new(threads, 32, 2, 0.0, True)
    orc.model_init(1, 0, 0.5, 2.0, 1)
    have_ref = ref_available("f64")
    if have_ref:
        ref = Ref("f64").init(threads, 32, 2, 0.0, True)
        ref.shift_init(2.0, 1)
    port_runs, ref_runs = [], []
    for _ in range(5):
        port_runs.append(sweep(lambda l, M: orc.sample(l, M)))
        if have_ref:
            ref_runs.append(sweep(lambda l, M: ref.shift_sample(l, M)))
    port, port_per = best(port_runs)
    out = {"cores": threads, "M0": M0, "levels": levels, "port": port, "port_per_level": port_per}
    if have_ref:
        rv, rper = best(ref_runs)
        out.update(reference=rv, reference_per_level=rper, port_over_reference=round(port / rv, 3),
                   ratio_per_run=[round(a / b, 3) for (a, _), (b, _) in zip(port_runs, ref_runs)])
    return out


def gpu_nonadaptive_sweep(nm, M0, levels, plain_cores):
    """Config 1's sweep on the GPU forms: per-level outer samples/s and the whole
    sweep's rate (non-adaptive: the reference rule with r = 0)."""
    out = {}
    for form, cores, label in ((nm.FORM_PLAIN, plain_cores, "plain_fp64"), (nm.FORM_BATCHED, 1, "batched_fp64"),
                               (nm.FORM_MIXED, 1, "mixed")):
        nm.init_nestedmc_sampler(cores, 32, 2, 0.0, True)
        nm.set_form(form)
        nm.model_init_ex(nm.MODEL_SHIFT, nm.PHI_INDICATOR, 0.5, 2.0, 1)
        sweep(nm, M0, levels)
        recs = []
        t0 = time.perf_counter()
        sweep(nm, M0, levels, recs)
        el = time.perf_counter() - t0
        total = sum(level_M(M0, l) for l in range(levels + 1))
        out[label] = {"outer_samples_per_s": round(total / el, 1), "per_level": per_level(recs, levels)}
        nm.cleanup_nestedmc_sampler()
    return out


def time_to_eps(nm, form, cores, eps, M0, setup_fn=None):
    (setup_fn or setup)(nm, form, cores, seed=7)
    t0 = time.perf_counter()
    r = nm.mlmc_run(eps, M0)
    dt = time.perf_counter() - t0
    Eg = r["levels"][r["starting_level"]]["sums_fine"][0] / r["levels"][r["starting_level"]]["M"]
    for lv in r["levels"][r["starting_level"] + 1:]:
        Eg += lv["sums_diff"][0] / lv["M"]
    rms = math.sqrt(r["bias_est"] ** 2 + r["stat_error"] ** 2)
    nm.cleanup_nestedmc_sampler()
    # Levels the driver froze: after the starting level, a level whose sampled
    # difference variance is zero gets no further samples (get_optimal_Ml), so
    # its mean enters the estimate as sampled from its first M0 draws only.
    frozen = [l for l, lv in enumerate(r["levels"]) if l > r["starting_level"] and lv["M"] > 0
              and lv["sums_diff"][1] / lv["M"] - (lv["sums_diff"][0] / lv["M"]) ** 2 <= 0.0]
    return {"seconds": round(dt, 3), "estimate": Eg, "rmse_est": rms, "levels": len(r["levels"]),
            "stat_error": r["stat_error"], "M": [lv["M"] for lv in r["levels"]],
            "zero_variance_levels": frozen}


def portfolio_configs(nm):
    import importlib
    return importlib.import_module(nm.__name__ + ".configs")


def book_sweep(nm, which, levels, M0, form, tune=True, cores=1):
    """Only a book's sweep (levels 0..levels, M_l = ceil(M0 2^-1.5l)):
    generation, launch-shape autotuning (as the portfolio leg does; default
    shapes with tune false), one warm-up sweep and one measured sweep in the
    given form, for profiler and counter runs.  Prints one JSON line."""
    cfg = portfolio_configs(nm)
    a, L, sc = cfg.large_portfolio() if which == "large" else cfg.small_portfolio()
    nm.init_nestedmc_sampler(cores, 64, 2, 0.0, True)
    nm.set_adaptive_rule(1, 4)
    nm.set_form(form)
    pf = nm.Portfolio(a["S0"], a["mu"], a["sigma"], a["counts"], a["strike"], a["maturity"], a["w"],
                      a["stock_ind"], a["type"], a["prob"], a["W_est"], **sc)
    pf.comp_params()
    nm.portfolio_init(pf, 1)
    nm.portfolio_set_correlation(L)
    shapes = autotune(nm, M0, levels) if (tune and form != nm.FORM_PLAIN) else None
    sweep(nm, M0, levels)
    recs = []
    t0 = time.perf_counter()
    sweep(nm, M0, levels, recs)
    el = time.perf_counter() - t0
    print(json.dumps({"book": which, "form": form, "levels": levels, "M0": M0, "seconds": round(el, 4),
                      "launch_shapes": shapes,
                      "kernel_ms": [round(x["kernel_ms"], 3) for x in recs],
                      "inner_samples": [x["inner"] for x in recs],
                      "rechecked_decisions": sum(x["rechecks"] for x in recs)}), flush=True)
    nm.cleanup_nestedmc_sampler()


def batch_tune(nm, n, levels, M0):
    """Launch-shape autotuning of a batch's level calls before the timed
    region, as the single-problem legs tune theirs: level l at M_l = ceil(M0
    2^-1.5l) outer samples per instance (the tuning calls draw outer samples
    set apart for tuning).  Returns [segment, blocks per CU] per level."""
    return [list(nm.batch_autotune_level(ell, [level_M(M0, ell)] * n)[:2]) for ell in range(levels + 1)]


def batch_sweep(nm, n, asian_frac, levels, M0, form, tune=True):
    """Only a level sweep of a batch of n books (configs.batch_books; with
    asian_frac > 0 the Asians on level-coupled paths): levels 0..levels with
    M_l = ceil(M0 2^-1.5l) outer samples per instance, launch shapes autotuned
    per level (tune), one warm-up and one measured sweep in the given form, for
    profiler and counter runs.  Prints one JSON line."""
    cfg = portfolio_configs(nm)
    books = cfg.batch_books(n, asian_frac=asian_frac)
    nm.init_nestedmc_sampler(1, 64, 2, 0.0, True)
    nm.set_adaptive_rule(1, 4)
    nm.set_form(form)
    nm.set_level_coupled_paths(asian_frac > 0)
    nm.batch_begin(n)
    keep = []
    for i, (a, L, sc, q) in enumerate(books):
        pf = nm.Portfolio(a["S0"], a["mu"], a["sigma"], a["counts"], a["strike"], a["maturity"], a["w"],
                          a["stock_ind"], a["type"], a["prob"], a["W_est"], **sc)
        pf.comp_params()
        keep.append(pf)
        nm.batch_stage(i, pf, 1000 + i, L)
    shapes = batch_tune(nm, n, levels, M0) if tune else None
    recs = []
    for rep in range(2):
        for ell in range(levels + 1):
            M = level_M(M0, ell)
            t0 = time.perf_counter()
            g = nm.batch_sample(ell, [M] * n)
            wall = time.perf_counter() - t0
            if rep:
                recs.append({"ell": ell, "M_per_instance": M, "kernel_ms": round(nm.last_kernel_ms(), 3),
                             "wall_ms": round(1e3 * wall, 3), "inner": sum(int(x["stats"][0]) for x in g)})
    print(json.dumps({"batch": n, "asian_frac": asian_frac, "coupled": asian_frac > 0, "form": form,
                      "launch_shapes": shapes, "levels": recs}), flush=True)
    nm.set_level_coupled_paths(False)
    nm.cleanup_nestedmc_sampler()


def portfolio_bench(nm, which, levels, M0, rel_eps, tte, plain_cores, tte_M0, tune=True, rel_eps2=None):
    """Portfolio books: per-level outer samples/s of the mixed and plain forms
    (N_base = 64 2^l, 3-sigma doubling capped at 2^4, M_l = ceil(M0 2^-1.5l)),
    their rooflines, the per-level comparison of the mixed form with the
    batched FP64 form on the same outer samples, and time to RMSE
    eps = rel_eps * (1 - q) (q the threshold's quantile), and with rel_eps2 at
    that tighter tolerance too."""
    cfg = portfolio_configs(nm)
    t0 = time.perf_counter()
    a, L, sc = cfg.large_portfolio() if which == "large" else cfg.small_portfolio()
    q = 0.995 if which == "large" else 0.99
    gen_s = time.perf_counter() - t0
    keep = {}

    def pf_setup(nm_, form, cores, seed=1):
        nm_.init_nestedmc_sampler(cores, 64, 2, 0.0, True)
        nm_.set_adaptive_rule(1, 4)
        nm_.set_form(form)
        pf = nm_.Portfolio(a["S0"], a["mu"], a["sigma"], a["counts"], a["strike"], a["maturity"],
                           a["w"], a["stock_ind"], a["type"], a["prob"], a["W_est"], **sc)
        pf.comp_params()
        nm_.portfolio_init(pf, seed)
        nm_.portfolio_set_correlation(L)
        keep["pf"] = pf

    out = {"workload": "%s book: d=%d assets, P=%d Exact options, threshold = pilot %.1f%% "
                       "quantile (%.4g)" % (which, len(a["S0"]), len(a["w"]), 100 * q,
                                            sc["max_loss"]),
           "levels": levels, "M0": M0, "generate_s": round(gen_s, 2)}
    eps = rel_eps * (1 - q)
    for form, cores, label in ((nm.FORM_MIXED, 1, "mixed"), (nm.FORM_PLAIN, plain_cores, "plain_fp64")):
        pf_setup(nm, form, cores)
        shapes = autotune(nm, M0, levels) if (tune and form != nm.FORM_PLAIN) else None
        sweep(nm, M0, levels)
        recs = []
        t1 = time.perf_counter()
        sweep(nm, M0, levels, recs)
        el = time.perf_counter() - t1
        total = sum(level_M(M0, l) for l in range(levels + 1))
        r = {"outer_samples_per_s": round(total / el, 1), "per_level": per_level(recs, levels),
             "kernel_ms": [round(x["kernel_ms"], 3) for x in recs],
             "inner_samples": sum(x["inner"] for x in recs),
             "roofline": portfolio_roofline(recs, "mixed" if form == nm.FORM_MIXED else "fp64")}
        if form == nm.FORM_MIXED:
            r["launch_shapes"] = shapes
            r["reevaluated"] = sum(x["reev"] for x in recs)
            r["flips"] = sum(x["flips"] for x in recs)
            r["rechecked_decisions"] = sum(x["rechecks"] for x in recs)
        nm.cleanup_nestedmc_sampler()
        if tte:
            r["time_to_eps"] = time_to_eps(nm, form, cores, eps, tte_M0, setup_fn=pf_setup)
            if rel_eps2:
                r["time_to_eps_tight"] = time_to_eps(nm, form, cores, rel_eps2 * (1 - q), tte_M0,
                                                     setup_fn=pf_setup)
        out[label] = r
        print("# portfolio %s %s: %.3e outer samples/s%s" % (
            which, label, r["outer_samples_per_s"],
            (", time to eps %.2f s" % r["time_to_eps"]["seconds"]) if tte else ""),
            file=sys.stderr, flush=True)
    out["mixed_vs_fp64_per_level"] = compare_forms(nm, pf_setup, M0, levels)
    if tte:
        for key, e in (("time_to_eps", eps), ("time_to_eps_tight", rel_eps2 * (1 - q) if rel_eps2 else None)):
            if e is None:
                continue
            m, p_ = out["mixed"][key], out["plain_fp64"][key]
            # The driver's stat_error is sqrt(Ca sum V/M) with Ca = 3: one
            # standard error is stat_error / sqrt(3).
            sig = math.hypot(m["stat_error"], p_["stat_error"]) / math.sqrt(3.0)
            out[key] = {"eps": e, "rel_eps": e / (1 - q), "speedup": round(p_["seconds"] / m["seconds"], 2),
                        "estimates_within_3_sigma": bool(abs(m["estimate"] - p_["estimate"]) <= 3 * sig)}
    return out


def batch_bench(nm, n, rel_eps, M0, plain_cores=0, asian_frac=0.0, coupled=False, tune=True):
    """Batched risk: n independent books (configs.batch_books: d in {10, 50,
    100}, P log-uniform in [1e3, 1e5], thresholds at the 95/99/99.5% pilot
    quantiles), each driven by its own MLMC run to eps_i = rel_eps (1 - q_i),
    all runs together over one batch of problems (one launch chain per level
    call for all instances): the whole batch's time to tolerance in the mixed
    and the batched FP64 forms, and whether each instance's two estimates agree
    within 3 combined standard errors; with plain_cores, also the plain FP64
    form run problem after problem (the reference's way: one problem per
    process).  With coupled, the Asians take level-coupled Milstein paths
    (dt_h0inv 2^floor(l/2) steps at level l, fine and coarse values from the
    same increments) instead of randomised-level paths.  With tune, each form's
    level calls 0..12 are autotuned first (batch_tune, M_l = ceil(2048
    2^-1.5l) per instance), outside the timed region."""
    cfg = portfolio_configs(nm)
    t0 = time.perf_counter()
    books = cfg.batch_books(n, asian_frac=asian_frac)
    gen_s = time.perf_counter() - t0
    tols = [rel_eps * (1 - q) for (_, _, _, q) in books]
    out = {"workload": "%d independent books: d in {10,50,100}, P log-uniform in [1e3,1e5] options "
                       "(%g%% arithmetic Asians by %s Milstein paths, 20%% digitals, the "
                       "rest calls/puts; t3 weights), factor-model correlation, thresholds at the pilot "
                       "95/99/99.5%% quantiles; per-instance eps = %g (1 - q)"
                       % (n, 100 * asian_frac, "level-coupled" if coupled else "randomised-level", rel_eps),
           "instances": n, "options_total": int(sum(len(b[0]["w"]) for b in books)),
           "generate_s": round(gen_s, 2)}
    runs = {}
    for form, label in ((nm.FORM_MIXED, "mixed"), (nm.FORM_BATCHED, "batched_fp64")):
        nm.init_nestedmc_sampler(1, 64, 2, 0.0, True)
        nm.set_adaptive_rule(1, 4)
        nm.set_form(form)
        nm.set_level_coupled_paths(coupled)
        nm.batch_begin(n)
        keep = []
        for i, (a, L, sc, q) in enumerate(books):
            pf = nm.Portfolio(a["S0"], a["mu"], a["sigma"], a["counts"], a["strike"], a["maturity"],
                              a["w"], a["stock_ind"], a["type"], a["prob"], a["W_est"], **sc)
            pf.comp_params()
            keep.append(pf)
            nm.batch_stage(i, pf, 1000 + i, L)
        shapes = batch_tune(nm, n, 12, 2048) if tune else None
        t1 = time.perf_counter()
        r = nm.batch_mlmc_run(tols, M0)
        dt = time.perf_counter() - t1
        nm.set_level_coupled_paths(False)
        nm.cleanup_nestedmc_sampler()
        runs[label] = r
        out[label] = {"seconds": round(dt, 3), "max_levels": max(len(x["levels"]) for x in r),
                      "outer_samples": int(sum(lv["M"] for x in r for lv in x["levels"])),
                      "launch_shapes": shapes}
        print("# batch %s: whole-batch time to tolerance %.2f s" % (label, dt), file=sys.stderr, flush=True)
    if plain_cores:
        # The reference's way: one problem at a time, plain FP64 form.
        t1 = time.perf_counter()
        for i, (a, L, sc, q) in enumerate(books):
            nm.init_nestedmc_sampler(plain_cores, 64, 2, 0.0, True)
            nm.set_adaptive_rule(1, 4)
            nm.set_form(nm.FORM_PLAIN)
            pf = nm.Portfolio(a["S0"], a["mu"], a["sigma"], a["counts"], a["strike"], a["maturity"],
                              a["w"], a["stock_ind"], a["type"], a["prob"], a["W_est"], **sc)
            pf.comp_params()
            nm.portfolio_init(pf, 1000 + i)
            nm.portfolio_set_correlation(L)
            nm.mlmc_run(tols[i], M0)
            nm.cleanup_nestedmc_sampler()
        dt = time.perf_counter() - t1
        out["plain_fp64_one_by_one"] = {"seconds": round(dt, 3), "virtual_threads": plain_cores}
        out["speedup_vs_plain_one_by_one"] = round(dt / out["mixed"]["seconds"], 2)
        print("# batch plain FP64 one problem at a time: %.2f s" % dt, file=sys.stderr, flush=True)
    em = [nm.mlmc_estimate(x) for x in runs["mixed"]]
    eb = [nm.mlmc_estimate(x) for x in runs["batched_fp64"]]
    within = [abs(a - b) <= 3 * math.hypot(x["stat_error"], y["stat_error"])
              for a, b, x, y in zip(em, eb, runs["mixed"], runs["batched_fp64"])]
    out["speedup_time_to_tolerance"] = round(out["batched_fp64"]["seconds"] / out["mixed"]["seconds"], 2)
    out["instances_within_3_sigma"] = int(sum(within))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-n", type=int, default=64, help="instances of the batched-risk leg (0: skip)")
    ap.add_argument("--batch-rel-eps", type=float, default=0.25)
    ap.add_argument("--batch-M0", type=int, default=1024)
    ap.add_argument("--batch-asian-frac", type=float, default=0.2,
                    help="share of Asian options in the books of the second batched-risk leg (0: skip it)")
    ap.add_argument("--batch-asian-rel-eps", type=float, default=0.25)
    ap.add_argument("--batch-asian-n", type=int, default=16,
                    help="instances of the Asian batched-risk leg (level-coupled paths)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the default launch shapes instead of autotuning each level first")
    ap.add_argument("--M0", type=float, default=1e7)
    ap.add_argument("--levels", type=int, default=12)
    ap.add_argument("--plain-cores", type=int, default=65536)
    ap.add_argument("--eps", type=float, default=1e-5)
    ap.add_argument("--tte-M0", type=int, default=1024)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tte", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-M0", type=float, default=1e6)
    ap.add_argument("--portfolio", choices=["large", "small", "none"], default="large")
    ap.add_argument("--pf-levels", type=int, default=6)
    ap.add_argument("--pf-M0", type=float, default=65536)
    ap.add_argument("--pf-rel-eps", type=float, default=0.05)
    ap.add_argument("--pf-tte-M0", type=int, default=1024)
    ap.add_argument("--pf-rel-eps2", type=float, default=0.025,
                    help="a second, tighter large-book tolerance (x (1 - q)); 0: skip")
    ap.add_argument("--cmp-M0", type=float, default=1e6,
                    help="M0 of the per-level mixed-vs-FP64 comparison sweep of the model")
    ap.add_argument("--nonadaptive-M0", type=float, default=1e6)
    ap.add_argument("--no-nonadaptive", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the small book")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (for scaling rocprofv3 byte counters)")
    ap.add_argument("--book-sweep-form", choices=["mixed", "plain"], default="mixed",
                    help="form of --book-sweep (plain: --plain-cores virtual threads)")
    ap.add_argument("--book-sweep", choices=["large", "small"],
                    help="only run one warm-up and one measured sweep of that book in --book-sweep-form "
                         "(autotuned launch shapes unless --no-autotune; for profiler and counter runs)")
    ap.add_argument("--batch-sweep", choices=["exact", "asian"],
                    help="only run one warm-up and one measured level sweep of --batch-sweep-n books "
                         "(Exact-only, or with 20%% Asians on level-coupled paths) in --batch-sweep-form; "
                         "for profiler and counter runs")
    ap.add_argument("--batch-sweep-n", type=int, default=16)
    ap.add_argument("--batch-sweep-levels", type=int, default=11)
    ap.add_argument("--batch-sweep-M0", type=float, default=2048)
    ap.add_argument("--batch-sweep-form", choices=["mixed", "batched"], default="mixed")
    args = ap.parse_args()
    if args.batch_sweep:
        import __graft_entry__
        nm = __graft_entry__.package()
        nm.load()
        batch_sweep(nm, args.batch_sweep_n, 0.2 if args.batch_sweep == "asian" else 0.0, args.batch_sweep_levels,
                    args.batch_sweep_M0, nm.FORM_MIXED if args.batch_sweep_form == "mixed" else nm.FORM_BATCHED,
                    tune=not args.no_autotune)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.book_sweep:
        import __graft_entry__
        nm = __graft_entry__.package()
        nm.load()
        plain = args.book_sweep_form == "plain"
        book_sweep(nm, args.book_sweep, args.pf_levels, args.pf_M0, nm.FORM_PLAIN if plain else nm.FORM_MIXED,
                   tune=not args.no_autotune, cores=args.plain_cores if plain else 1)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the outer samples "
              "would shard with one all-reduce of ~300 B per level call); run with --gpus 1.")
        sys.exit(2)

    import __graft_entry__
    nm = __graft_entry__.package()
    nm.load()
    chosen = nm.FORM_MIXED

    # Chosen form: launch-shape autotuning, warmup, then K timed sweeps.
    setup(nm, chosen, 1)
    shapes = None if args.no_autotune else autotune(nm, args.M0, args.levels)
    for _ in range(args.warmup):
        sweep(nm, args.M0, args.levels)
    recs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sweep(nm, args.M0, args.levels, recs)
    elapsed = time.perf_counter() - t0
    outer_per_step = sum(level_M(args.M0, l) for l in range(args.levels + 1))
    value = outer_per_step * args.steps / elapsed
    reev = sum(r["reev"] for r in recs)
    flips = sum(r["flips"] for r in recs)
    rechecks = sum(r["rechecks"] for r in recs)
    rl = roofline(recs, chosen)
    levels_mixed = per_level(recs, args.levels)
    print("# mixed: %.3e outer samples/s, %.1f ms/step, roofline %.3f of VALU peak"
          % (value, 1e3 * elapsed / args.steps, rl["frac"]), file=sys.stderr, flush=True)
    nm.cleanup_nestedmc_sampler()

    extra = {}
    if not args.no_plain:
        setup(nm, nm.FORM_PLAIN, args.plain_cores)
        sweep(nm, args.M0, args.levels)
        precs = []
        t0 = time.perf_counter()
        sweep(nm, args.M0, args.levels, precs)
        pel = time.perf_counter() - t0
        extra["plain_fp64"] = {"outer_samples_per_s": round(outer_per_step / pel, 1),
                               "per_level": per_level(precs, args.levels),
                               "roofline_frac": roofline(precs, nm.FORM_PLAIN)["frac"],
                               "virtual_threads": args.plain_cores}
        print("# plain: %.3e outer samples/s" % (outer_per_step / pel), file=sys.stderr, flush=True)
        nm.cleanup_nestedmc_sampler()
        setup(nm, nm.FORM_BATCHED, 1)
        bshapes = None if args.no_autotune else autotune(nm, args.M0, args.levels)
        sweep(nm, args.M0, args.levels)
        brecs = []
        t0 = time.perf_counter()
        sweep(nm, args.M0, args.levels, brecs)
        bel = time.perf_counter() - t0
        extra["batched_fp64"] = {"outer_samples_per_s": round(outer_per_step / bel, 1),
                                 "per_level": per_level(brecs, args.levels), "launch_shapes": bshapes}
        nm.cleanup_nestedmc_sampler()
    extra["mixed_vs_fp64_per_level"] = compare_forms(nm, setup, args.cmp_M0, args.levels)
    if not args.no_tte:
        tm = time_to_eps(nm, chosen, 1, args.eps, args.tte_M0)
        print("# time to eps (mixed): %s" % tm, file=sys.stderr, flush=True)
        tp = time_to_eps(nm, nm.FORM_PLAIN, args.plain_cores, args.eps, args.tte_M0)
        print("# time to eps (plain): %s" % tp, file=sys.stderr, flush=True)
        exact = 0.5 * math.erfc(2 / math.sqrt(2))
        # stat_error = sqrt(Ca sum V/M) with Ca = 3: one standard error is stat_error / sqrt(3).
        sig = math.hypot(tm["stat_error"], tp["stat_error"]) / math.sqrt(3.0)
        extra["time_to_eps"] = {"eps": args.eps, "mixed": tm, "plain_fp64": tp,
                                "speedup": round(tp["seconds"] / tm["seconds"], 2),
                                "exact": exact,
                                "estimates_within_3_sigma": bool(abs(tm["estimate"] - tp["estimate"]) <= 3 * sig),
                                "mixed_error_over_rmse": abs(tm["estimate"] - exact) / tm["rmse_est"],
                                "plain_error_over_rmse": abs(tp["estimate"] - exact) / tp["rmse_est"],
                                "note": "a level listed in zero_variance_levels drew no nonzero difference "
                                        "in its first M0 outer samples and is frozen there by the driver's "
                                        "allocation (get_optimal_Ml, mlmc.hpp:265-289); its mean (level 3: "
                                        "-2.48e-5) is then missing from the estimate"}
    if args.portfolio != "none":
        extra["portfolio"] = portfolio_bench(nm, args.portfolio, args.pf_levels, args.pf_M0,
                                             args.pf_rel_eps, not args.no_tte, args.plain_cores,
                                             args.pf_tte_M0, tune=not args.no_autotune,
                                             rel_eps2=args.pf_rel_eps2 or None)
    if not args.no_small:
        extra["small_book"] = portfolio_bench(nm, "small", args.pf_levels, args.pf_M0, args.pf_rel_eps,
                                                      not args.no_tte, args.plain_cores, args.pf_tte_M0,
                                                      tune=not args.no_autotune)
    if not args.no_nonadaptive:
        extra["nonadaptive_sweep"] = {"workload": "shift model, adaptive refinement off (N = 32 2^l), antithetic, seed 1, "
                                        "levels 0..8, M_l = ceil(%g 2^-1.5l)" % args.nonadaptive_M0,
                            "gpu": gpu_nonadaptive_sweep(nm, args.nonadaptive_M0, 8, args.plain_cores)}
        if not args.no_cpu:
            extra["nonadaptive_sweep"]["cpu"] = cpu_nonadaptive_sweep(args.nonadaptive_M0)
    if args.batch_n > 0 and not args.no_tte:
        # The batch drivers' level calls keep the default launch shapes: a shape
        # tuned at one M per instance holds near that M, and the drivers' calls
        # range far from it (tuned at the sweep's M, the 64-book batch took
        # 5.26 s instead of 4.55 s and the Asian batch 140 s instead of 74 s).
        extra["batch"] = batch_bench(nm, args.batch_n, args.batch_rel_eps, args.batch_M0,
                                     0 if args.no_plain else args.plain_cores, tune=False)
        if args.batch_asian_frac > 0:
            extra["batch_asian"] = batch_bench(nm, args.batch_asian_n, args.batch_asian_rel_eps, args.batch_M0, 0,
                                               args.batch_asian_frac, coupled=True, tune=False)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_M0, args.levels)
        c1 = extra.get("nonadaptive_sweep", {}).get("cpu")
        if c1 and "reference" in c1:
            cpu["calibration"] = {"workload": "non-adaptive shift-model sweep, levels 0..8, same threads",
                                  "cores": c1["cores"], "port_outer_samples_per_s": c1["port"],
                                  "reference_outer_samples_per_s": c1["reference"],
                                  "port_over_reference": c1["port_over_reference"],
                                  "ratio_per_run": c1["ratio_per_run"]}

    line = {
        "metric": "outer samples/s per level; time to RMSE eps: batched FP32+FP64 vs plain FP64",
        "value": round(value, 1),
        "unit": "outer samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (counter-based Threefry draws generated on the device)",
        "config": {"workload": "adaptive model problem: X=Y+Z, threshold 2, levels 0..%d, "
                               "N_base=32*2^l, 3-sigma doubling capped at 2^6, "
                               "M_l=ceil(%g*2^-1.5l)" % (args.levels, args.M0),
                   "form": "batched FP32+FP64 (mixed)", "outer_samples_per_step": outer_per_step,
                   "launch_shapes": shapes},
        "per_level_outer_samples_per_s": levels_mixed,
        "mixed_reevaluated": reev,
        "mixed_flips": flips,
        "mixed_rechecked_decisions": rechecks,
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
