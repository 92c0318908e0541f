"""The oracle (C restatement) against the reference's own outputs.

Fixtures in tests/golden/ were written by tests/golden/make_goldens.py from the
reference's OpenMP build (compiled in place into oracle/_ref/).  Where that
build is present, the same comparisons also run live against it.  Bit-exact
for single-thread runs (same compiler family, same libm, unfused FP64); sums of
multi-thread runs agree to 1e-14 (the reference adds thread results in
arbitrary order).
"""
import json
import os

import numpy as np
import pytest

from oracle import Oracle, Ref, ref_available

HERE = os.path.dirname(os.path.abspath(__file__))
REALS = ["f64", "f32"]


def fixture(real):
    with open(os.path.join(HERE, "golden", "reference_%s.json" % real)) as f:
        return json.load(f)


def u64(a):
    return np.array([int(x) for x in a], dtype=np.uint64)


@pytest.mark.parametrize("real", REALS)
def test_threefry_and_streams(real):
    fx = fixture(real)["rng"]
    O = Oracle(real)
    for c in fx["threefry"]:
        assert O.threefry([int(x) for x in c["ctr"]], [int(x) for x in c["key"]]) == \
            [int(x) for x in c["out"]]
    for s in fx["streams"]:
        u, r = O.stream(s["seed"], s["core"], s["kinds"])
        assert np.array_equal(u, u64(s["u"]))
        assert np.array_equal(r, np.array(s["r"]))  # exact, including the randn cache
    for v in fx["u01"]:
        x = int(v["u"])
        assert O.u01_f64(x) == v["u01_f64"]
        assert O.uneg11_f64(x) == v["uneg11_f64"]
        assert np.float32(O.u01_f32(x)) == np.float32(v["u01_f32"])
    for b in fx["boxmuller"]:
        assert list(O.boxmuller(int(b["u0"]), int(b["u1"]))) == b["out"]


@pytest.mark.parametrize("real", REALS)
def test_utilities(real):
    fx = fixture(real)["util"]
    O = Oracle(real)
    for c in fx["cdf"]:
        assert [O.rand_int_cdf(u, c["cdf"]) for u in c["u"]] == c["idx"]
    for c in fx["pdf"]:
        assert [O.rand_int_pdf(u, c["pdf"]) for u in c["u"]] == c["idx"]
    for p, q in fx["norm_cdf_inv"]:
        assert O.norm_cdf_inv(p) == q
    for x, q in fx["norm_cdf"]:
        assert O.norm_cdf(x) == q
    for b in fx["bs"]:
        assert O.bs_price(b["type"], *b["args"]) == b["price"]
        assert O.bs_delta(b["type"], *b["args"]) == b["delta"]
    for c in fx["ssdt"]:
        W, S = O.simulate_stock_dt(c["with_coarse"], c["normals"], c["S"], c["N"], 0.05, 0.3, 0.75,
                                   3, c["milstein"], c["coarse_step"])
        assert W == c["W"]
        assert np.array_equal(S, np.array(c["S_out"]))


def _cmp_call(o, g, exact):
    assert np.array_equal(o["stats"], u64(g["stats"]))
    assert o["work"] == g["work"]
    if exact:
        assert np.array_equal(o["fsums"], np.array(g["fsums"]))
        assert np.array_equal(o["diffsums"], np.array(g["diffsums"]))
    else:
        np.testing.assert_allclose(o["fsums"], g["fsums"], rtol=1e-14, atol=0)
        np.testing.assert_allclose(o["diffsums"], g["diffsums"], rtol=1e-14, atol=1e-300)


@pytest.mark.parametrize("real", REALS)
def test_model_level_calls(real):
    """Consecutive model_sample calls: streams persist across calls and levels."""
    O = Oracle(real)
    for run in fixture(real)["model"]:
        p = run["params"]
        O.new(p["cores"], p["N0"], p["beta"], p["r"], p["anti"])
        O.model_init(0, p["phi"], run["tau"], run["max_loss"], p["seed"])
        for g in run["calls"]:
            _cmp_call(O.sample(g["ell"], g["M"]), g, exact=p["cores"] == 1)


@pytest.mark.parametrize("real", REALS)
def test_model_per_sample(real):
    """One outer sample per call (M=1, one thread): per-sample phi_f, phi_f - phi_c."""
    fx = fixture(real)["model_per_sample"]
    p = fx["params"]
    O = Oracle(real).new(1, p["N0"], p["beta"], p["r"], p["anti"])
    O.model_init(0, p["phi"], fx["tau"], fx["max_loss"], p["seed"])
    for s in fx["samples"]:
        o = O.sample(s["ell"], 1)
        assert o["fsums"][0] == s["phi_f"]
        assert o["diffsums"][0] == s["dphi"]
        assert np.array_equal(o["stats"], u64(s["stats"]))


def _portfolio(gen, scalars, start, delta):
    pf = {k: np.array(v) for k, v in gen.items()}
    pf["start"], pf["delta_est"] = np.array(start), np.array(delta)
    pf.update(scalars)
    return pf


@pytest.mark.parametrize("real", REALS)
def test_portfolio_level_calls(real):
    """Portfolio problem: BS, Exact and Unbiased (Milstein) options, every
    sub-sampling mode and phi type, on the reference's generated portfolio."""
    fx = fixture(real)["portfolio"]
    O = Oracle(real)
    gen = fx["gen"]["data"]
    for run in fx["runs"]:
        O.new(run["cores"], run["N0"], run["beta"], run["r"], True)
        pf = _portfolio(gen, run["scalars"], run["start"], run["delta_est"])
        # portfolio_comp_params: start / delta_est are the BS value and delta.
        for j, (typ, K, T, si) in enumerate(zip(gen["type"], gen["strike"], gen["maturity"],
                                               gen["stock_ind"])):
            S0, sig = gen["S0"][si], gen["sigma"][si]
            assert O.bs_price(typ, S0, K, pf["r"], T, sig) == run["start"][j]
            assert O.bs_delta(typ, S0, K, pf["r"], T, sig) == run["delta_est"][j]
        O.portfolio_init(pf, 3)
        for g in run["calls"]:
            _cmp_call(O.sample(g["ell"], g["M"]), g, exact=run["cores"] == 1)


@pytest.mark.parametrize("real", REALS)
def test_portfolio_delta_unbiased(real):
    """portfolio_sample_delta_unbiased: moments and work of consecutive calls
    (streams persist) under 1 and 3 threads, Milstein paths of the generated
    portfolio's Unbiased options."""
    fx = fixture(real)["portfolio"]
    O = Oracle(real)
    gen = fx["gen"]["data"]
    run0 = fx["runs"][0]
    for case in fx["delta_unbiased"]:
        O.new(case["cores"], 8, 4, 1.5, True)
        scal = dict(run0["scalars"], subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                    phi_type=0, phi_size=1)
        O.portfolio_init(_portfolio(gen, scal, run0["start"], run0["delta_est"]), case["seed"])
        for g in case["calls"]:
            sums, work = O.sample_delta_unbiased(g["option_index"], g["M"], 2)
            assert work == g["work"]
            if case["cores"] == 1:
                assert sums.tolist() == g["sums"]
            else:
                np.testing.assert_allclose(sums, g["sums"], rtol=1e-14, atol=1e-300)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("real", REALS)
def test_live_against_reference_build(real):
    """Random parameter draws, live against the reference build.  beta is 4 or 8:
    with beta = 2 and the adaptive rule the reference evaluates uint_pow(1, l),
    whose clog2(1) = 64 - clz(0) is undefined (the oracle and the library return
    1 there), so its output is not a reference value."""
    O, R = Oracle(real), Ref(real)
    rs = np.random.RandomState(5)
    for _ in range(6):
        cores = int(rs.choice([1, 2, 5]))
        N0 = int(rs.choice([4, 8, 32]))
        beta = int(rs.choice([4, 8]))
        r = float(rs.choice([0.0, 1.3, 1.7]))
        anti = bool(rs.randint(2))
        phi = int(rs.randint(2))
        seed = int(rs.randint(1000))
        ml = float(rs.uniform(-0.02, 0.03))
        O.new(cores, N0, beta, r, anti)
        R.init(cores, N0, beta, r, anti)
        O.model_init(0, phi, 0.02, ml, seed)
        R.model_init(phi, 0.02, ml, seed)
        for ell in (0, 2, 1, 3):
            M = int(rs.randint(1, 40))
            _cmp_call(O.sample(ell, M), {k: (v.tolist() if hasattr(v, "tolist") else v)
                                         for k, v in R.model_sample(ell, M).items()},
                      exact=cores == 1)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
def test_shift_model_against_reference_engine():
    """The benchmark's shift model (X = Y + Z, threshold 2, Indicator) as a
    sampler class on the reference's own engine, against the oracle's shift
    model under the reference keying: the non-adaptive sweep of levels 0..8
    with N = 32 2^l, antithetic, seed 1, 8 threads (the CPU-path workload at a
    reduced M0).  Work and counts exact; sums to 1e-12 (the reference adds its
    threads' sums in completion order)."""
    O, R = Oracle("f64"), Ref("f64")
    O.new(8, 32, 2, 0.0, True)
    R.init(8, 32, 2, 0.0, True)
    O.model_init(1, 0, 0.5, 2.0, 1)
    R.shift_init(2.0, 1)
    for ell in range(9):
        M = int(np.ceil(4000 * 2.0 ** (-1.5 * ell)))
        o, r = O.sample(ell, M), R.shift_sample(ell, M)
        assert o["work"] == r["work"] and np.array_equal(o["stats"], r["stats"])
        np.testing.assert_allclose(o["fsums"], r["fsums"], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(o["diffsums"], r["diffsums"], rtol=1e-12, atol=1e-300)
