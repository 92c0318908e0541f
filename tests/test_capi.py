"""C-ABI of libnmlmc.so on the CPU: every declared entry point is exported, the
status-returning paths report errors without touching a GPU, the struct layouts
match the reference's REAL_DOUBLE build, and the host MLMC driver
(mlmc_t::run) reproduces the reference's call sequence on deterministic level
samplers."""
import ctypes as C
import json
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            name = m.group(1)
            if name not in ("if", "sizeof") and not text[m.start():m.end()].lstrip().startswith("typedef"):
                names.append(name)
    return sorted(set(names))


def test_header_symbols_exported(nm):
    lib = nm.load()
    names = declared_functions()
    assert "model_sample" in names and "nmlmc_mlmc_run" in names and len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(nm.EXPORTS) <= set(names)


def test_struct_layouts_match_reference_real_double(nm):
    """sizeof / offsetof of the ABI structs equal the reference's -DREAL_DOUBLE
    build (probed from the compiled reference into the golden fixture)."""
    with open(os.path.join(ROOT, "tests", "golden", "reference_f64.json")) as f:
        ref = json.load(f)["layout"]
    P, S, D, O, K, F = (nm.NestedSamplerParams, nm.SampleOutput, nm.ModelData, nm.OptionParams,
                        nm.StockParams, nm.PortfolioData)
    mine = [C.sizeof(P), P.beta.offset, P.antithetic.offset, P.fine_moments.offset,
            P.diff_moments.offset, P.adaptive_r.offset,
            C.sizeof(S), S.fsums.offset, S.diffsums.offset, S.total_work.offset, S.inner_stats.offset,
            C.sizeof(D), D.s1tau.offset, D.max_loss.offset, D.phi_type.offset,
            C.sizeof(O), O.maturity.offset, O.start.offset, O.w.offset, O.delta_est.offset,
            O.stock_ind.offset, O.type.offset, O.prob.offset, O.W_est.offset,
            C.sizeof(K),
            C.sizeof(F), F.stocks.offset, F.r.offset, F.rho.offset, F.s1_rho2.offset, F.max_loss.offset,
            F.tau.offset, F.dt_h0inv.offset, F.dt_beta.offset, F.dt_prob_expo.offset, F.phi_type.offset,
            F._phi_size.offset, F.subsample.offset, F.use_delta_cv.offset, F.use_coarse_delta_cv.offset,
            F.use_antithetic_risk.offset, F.use_milstein.offset]
    assert mine == ref


def test_errors_without_a_gpu(nm):
    lib = nm.load()
    out = nm.Output()
    rc = lib.nmlmc_sample(0, 10, (1 << 64) - 1, C.byref(out.c), None, None, None, None)
    assert rc == 1 and lib.nmlmc_last_error() == 1  # NMLMC_ERR_STATE: nothing initialised
    assert b"init_nestedmc_sampler" in lib.nmlmc_last_error_string()
    assert lib.nmlmc_set_form(7) == 2  # NMLMC_ERR_ARG
    assert lib.nmlmc_set_adaptive_rule(1, 3) == 1  # needs init first
    lib.init_nestedmc_sampler(0, None)
    assert lib.nmlmc_last_error() == 2
    # Batch calls check their array length against the staged batch (none here).
    Ms = (C.c_uint64 * 2)(10, 10)
    outs = (nm.SampleOutput * 2)()
    assert lib.nmlmc_batch_sample(0, 2, Ms, outs) == 2
    assert b"batch of 0" in lib.nmlmc_last_error_string()
    with pytest.raises(nm.NmlmcError):
        nm.batch_sample(0, [10, 10])


def test_library_matches_sources(nm):
    """The loaded library was built from the sources beside it."""
    assert nm.load().nmlmc_source_digest().decode() == nm.source_digest()


def _two_point(mean, var, M, real=float):
    s = math.sqrt(var)
    npos, nneg = (M + 1) // 2, M // 2
    a, b = mean + s, mean - s
    return [npos * math.pow(a, k + 1) + nneg * math.pow(b, k + 1) for k in range(4)]


def test_mlmc_driver_reproduces_reference_call_sequence(nm):
    """The host driver against the reference's mlmc_t::run (mlmc.hpp:290-337)
    on the same deterministic level sampler: identical (ell, M) calls."""
    with open(os.path.join(ROOT, "tests", "golden", "reference_f64.json")) as f:
        fx = json.load(f)["mlmc"]
    for case in fx:
        cs = case["case"]
        calls = []
        L = len(cs["E"])

        def level(ell, M):
            l = min(ell, L - 1)
            calls.append((ell, M))
            return dict(sums_fine=_two_point(cs["Ef"][l], cs["Vf"][l], M),
                        sums_diff=_two_point(cs["E"][l], cs["V"][l], M),
                        work=cs["W"][l] * M, M=M)

        r = nm.mlmc_run_cb(level, cs["tol"], cs["M0"])
        assert [c[0] for c in calls] == case["ell"], cs["name"]
        assert [c[1] for c in calls] == [int(x) for x in case["M"]], cs["name"]
        assert len(r["levels"]) == case["summary"][0]
        assert r["starting_level"] == case["summary"][1]
        assert r["bias_est"] == pytest.approx(case["summary"][2], rel=1e-15)
        assert r["stat_error"] == pytest.approx(case["summary"][3], rel=1e-15)
        assert r["TOL"] == pytest.approx(case["summary"][4], rel=1e-15)
        # The level table (mlmc_output_t::to_string, mlmc.hpp:106-183), the
        # format parse_mlmc_output.py reads: identical text.
        assert nm.mlmc_last_table() == case["table"], cs["name"]


def test_mlmc_driver_closed_form_allocation(nm):
    """Closed form of get_optimal_Ml: with V_l = 4^-l, W_l = 2^l and E_l = 2^-l
    the driver's final sample counts satisfy M_l >= lambda sqrt(V_l / W_l)."""
    L = 14

    def level(ell, M):
        return dict(sums_fine=_two_point(0.5, 0.25, M),
                    sums_diff=_two_point(2.0 ** -(ell + 1), 4.0 ** -(ell + 1), M),
                    work=2.0 ** ell * M, M=M)

    r = nm.mlmc_run_cb(level, 1e-3, 100)
    lv = r["levels"]
    assert r["stat_error"] ** 2 < 0.5 * r["TOL"] ** 2 * 2  # theta TOL^2 at the last TOL
    ratios = [lv[l]["M"] / lv[l + 1]["M"] for l in range(1, len(lv) - 1)]
    # M_l / M_{l+1} = sqrt(V_l W_{l+1} / (V_{l+1} W_l)) = sqrt(4 * 2) up to ceil().
    for q in ratios:
        assert q == pytest.approx(math.sqrt(8), rel=0.05)


def test_mlmc_driver_callback_error_aborts(nm):
    def bad(ell, M):
        raise RuntimeError("boom")
    with pytest.raises(nm.NmlmcError):
        nm.mlmc_run_cb(bad, 1e-2, 10)


def test_portfolio_gen_data_matches_reference(nm):
    """portfolio_gen_data (host only) against the reference build's output for
    the same arguments: every field bit-identical, including the order of
    equal sub-sampling probabilities, which std::sort decides."""
    with open(os.path.join(ROOT, "tests", "golden", "reference_f64.json")) as f:
        fx = json.load(f)["portfolio"]
    g0 = fx["gen"]
    cases = [dict(P=g0["P"], Q=g0["Q"], seed=g0["seed"], probs=g0["probs"], uniform_prob=False,
                  data=g0["data"])] + fx["gen_more"]
    for c in cases:
        arr, scal = nm.portfolio_gen_data(c["P"], c["Q"], c["seed"], c["probs"], False,
                                          c["uniform_prob"])
        for k, v in c["data"].items():
            assert np.array_equal(np.array(v, dtype=arr[k].dtype), arr[k]), (c["P"], k)
        assert scal["max_loss"] == 0.7 and scal["subsample"] == 0 and scal["use_milstein"]


def test_portfolio_gen_data_rejects_bad_split(nm):
    with pytest.raises(nm.NmlmcError):
        nm.portfolio_gen_data(10, 6, 1, [0.3, 0.5, 0.2])   # P < 2Q
    with pytest.raises(nm.NmlmcError):
        nm.portfolio_gen_data(10, 2, 1, [0.33, 0.33, 0.33])  # floors do not add up to P
