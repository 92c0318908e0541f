"""GPU parity of the model problem: HIP library vs the oracle on the same seeded
inputs (per outer sample), plus the closed-form answer of the shift model.

Tolerances: counts, work and inner_stats are integers and must match exactly;
phi values and sums are FP64 and must agree to 1e-12 relative (the GPU uses the
device libm for log/sincos and a parallel summation order)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _cmp(g, o, trace=True):
    assert np.array_equal(g["stats"], o["stats"]), (g["stats"], o["stats"])
    assert g["work"] == o["work"]
    np.testing.assert_allclose(g["fsums"], o["fsums"], rtol=RTOL, atol=1e-300)
    np.testing.assert_allclose(g["diffsums"], o["diffsums"], rtol=RTOL, atol=1e-300)
    if trace:
        assert np.array_equal(g["nf"], o["nf"])
        assert np.array_equal(g["nc"], o["nc"])
        np.testing.assert_allclose(g["phi_f"], o["phi_f"], rtol=RTOL, atol=1e-300)
        np.testing.assert_allclose(g["phi_c"], o["phi_c"], rtol=RTOL, atol=1e-300)


def _ref_max_loss(orc, tau=0.02, eta=0.025):
    # model_sampler.cpp:144: max_loss = tau (Phi^-1(eta/2)^2 - 1)
    return tau * (orc.norm_cdf_inv(eta / 2) ** 2 - 1)


@pytest.mark.parametrize("cores,adaptive_r,anti,phi", [
    (1, 0.0, True, 0), (8, 1.5, True, 0), (64, 1.5, True, 1), (64, 0.0, False, 1),
    (32, 1.5, False, 0), (3, 1.5, True, 1)])
def test_plain_reference_model_matches_oracle(nm, oracle_f64, cores, adaptive_r, anti, phi):
    """Plain FP64 form with the reference keying, reference model
    (model_sampler.cpp:42-92): per-sample phi_f, phi_c, N_f, N_c and level sums
    over consecutive calls (stream state persists across calls and levels)."""
    print("case", cores, adaptive_r, anti, phi, flush=True)
    N0, beta, tau = 32, 4, 0.02
    orc = oracle_f64.new(cores, N0, beta, adaptive_r, anti)
    ml = _ref_max_loss(orc, tau)
    nm.init_nestedmc_sampler(cores, N0, beta, adaptive_r, anti)
    nm.model_init(phi, tau, ml, 11)
    orc.model_init(0, phi, tau, ml, 11)
    for ell, M in ((0, 300), (1, 200), (3, 77), (2, 130), (0, 5)):
        _cmp(nm.sample(ell, M, trace=True), orc.sample(ell, M, trace=True))
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("lanes", [1, 4, 16, 64, 256])
def test_plain_lane_groups_do_not_change_results(nm, oracle_f64, lanes, monkeypatch):
    """The lanes-per-virtual-thread split is invisible in the results (256: one
    virtual thread per four-wave block, its sums reduced across the waves)."""
    monkeypatch.setenv("NMLMC_GROUP_LANES", str(lanes))
    orc = oracle_f64.new(16, 32, 4, 1.5, True)
    ml = _ref_max_loss(orc)
    nm.init_nestedmc_sampler(16, 32, 4, 1.5, True)
    nm.model_init(0, 0.02, ml, 3)
    orc.model_init(0, 0, 0.02, ml, 3)
    for ell, M in ((0, 100), (2, 60)):
        _cmp(nm.sample(ell, M, trace=True), orc.sample(ell, M, trace=True))
        assert nm.last_group_lanes() == lanes
    nm.cleanup_nestedmc_sampler()


def test_drop_in_model_sample_entry(nm, oracle_f64):
    """model_sample (model_sampler.cpp:106-114) through the reference-named entry."""
    orc = oracle_f64.new(8, 32, 4, 1.5, True)
    ml = _ref_max_loss(orc)
    nm.init_nestedmc_sampler(8, 32, 4, 1.5, True)
    nm.model_init(0, 0.02, ml, 0)
    orc.model_init(0, 0, 0.02, ml, 0)
    for ell, M in ((0, 64), (1, 33)):
        g = nm.model_sample(ell, M)
        o = orc.sample(ell, M)
        assert g["M"] == M
        _cmp(g, o, trace=False)
    nm.cleanup_nestedmc_sampler()


def _shift_setup(nm, orc, cores, form, seed=1, cap=6):
    nm.init_nestedmc_sampler(cores, 32, 2, 0.0, True)
    nm.set_adaptive_rule(1, cap)
    nm.set_form(form)
    nm.model_init_ex(nm.MODEL_SHIFT, 0, 0.5, 2.0, seed)
    o = orc.new(cores, 32, 2, 0.0, True, adapt_rule=1, cap_log2=cap)
    o.model_init(1, 0, 0.5, 2.0, seed)
    return o


def test_plain_shift_model_capped_rule(nm, oracle_f64):
    """Shift model X = Y + Z, threshold 2, 3-sigma doubling capped at 2^6."""
    orc = _shift_setup(nm, oracle_f64, 32, nm.FORM_PLAIN)
    for ell, M in ((0, 400), (1, 300), (4, 60)):
        _cmp(nm.sample(ell, M, trace=True), orc.sample(ell, M, trace=True))
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("segment", [0, 1, 3, 16, 64])
def test_batched_fp64_matches_oracle(nm, oracle_f64, segment, monkeypatch):
    """Batched FP64 form: per-outer-sample keys; identical per-sample results to
    the oracle's batched keying for any segment size (Box-Muller pairs per lane
    per work item), i.e. however the draw groups are split across lanes."""
    if segment:
        monkeypatch.setenv("NMLMC_SEGMENT", str(segment))
    orc = _shift_setup(nm, oracle_f64, 1, nm.FORM_BATCHED)
    for ell, M, off in ((0, 500, 0), (1, 300, 0), (3, 100, 0), (1, 200, 300)):
        g = nm.sample(ell, M, outer_offset=off, trace=True)
        o = orc.sample(ell, M, keying=1, outer_offset=off, trace=True)
        _cmp(g, o)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_autotuned_launch_shape_keeps_results(nm, oracle_f64):
    """The launch-shape autotuner times its candidates on outer samples set
    apart for tuning: the level calls after it draw the outer samples they would
    have drawn without it, and per-sample results equal the oracle's whatever
    shape it chose.  The plain form has nothing to tune."""
    orc = _shift_setup(nm, oracle_f64, 1, nm.FORM_BATCHED)
    for ell, M in ((0, 3000), (2, 400)):
        S, B, ms = nm.autotune_level(ell, M)
        assert S in (4, 8, 16, 32, 64, 128) and B in (4, 8, 16) and ms > 0
        g = nm.sample(ell, M, trace=True)
        assert nm.last_group_lanes() == S
        o = orc.sample(ell, M, keying=1, outer_offset=0, trace=True)
        _cmp(g, o)
    nm.set_form(nm.FORM_PLAIN)
    with pytest.raises(nm.NmlmcError):
        nm.autotune_level(0, 100)
    nm.cleanup_nestedmc_sampler()


def _sums(r):
    return (tuple(r["fsums"]), tuple(r["diffsums"]), r["work"], tuple(int(x) for x in r["stats"]), r["M"])


@pytest.mark.parametrize("graph", ["1", "0"])
@pytest.mark.parametrize("form", [1, 2])
def test_level_sweep_matches_lone_calls(nm, oracle_f64, form, graph, monkeypatch):
    """A level sweep (captured as one HIP graph, or enqueued directly) gives
    every level call the results of the lone call on the same outer samples.
    Three sweeps with eager calls and waits in between: each sweep (a fresh
    capture updating the executable graph) equals its lone calls, and levels
    continue their outer ids across sweeps and lone calls.  Level 0 runs as
    three chunks.  Indicator sums, work and counts are integers, so equal."""
    monkeypatch.setenv("NMLMC_SWEEP_GRAPH", graph)
    monkeypatch.setenv("NMLMC_BATCH_CHUNK", "1000")
    ells, Ms = [0, 1, 2, 3], [2500, 900, 300, 100]
    _shift_setup(nm, oracle_f64, 1, form)
    ref = []
    for _ in range(3):
        ref.append([_sums(nm.sample(l, M)) for l, M in zip(ells, Ms)] + [_sums(nm.sample(5, 40))])
    nm.cleanup_nestedmc_sampler()
    _shift_setup(nm, oracle_f64, 1, form)
    for rep in range(3):
        got = [_sums(r) for r in nm.sample_levels(ells, Ms)]
        got.append(_sums(nm.sample(5, 40)))
        assert got == ref[rep], rep
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("form", [1, 2])
def test_final_pool_overflow_redo(nm, oracle_f64, form, monkeypatch):
    """A level call whose final-group pool overflows grows the pool and redoes
    the call from the same keyed draws: a lone call in place, a call of a
    level sweep on its own after the sweep.  The results equal those of calls
    whose pool was large enough.  Lone calls and sweeps alternate, so the
    clearing of the control block between chains is crossed by the redos."""
    ells, Ms = [0, 1, 2, 3], [2500, 900, 300, 100]
    _shift_setup(nm, oracle_f64, 1, form)
    ref = [_sums(nm.sample(l, M)) for l, M in zip(ells, Ms)]
    ref2 = [_sums(nm.sample(l, M)) for l, M in zip(ells, Ms)]
    nm.cleanup_nestedmc_sampler()
    monkeypatch.setenv("NMLMC_FINAL_POOL", "64")
    _shift_setup(nm, oracle_f64, 1, form)
    assert [_sums(nm.sample(l, M)) for l, M in zip(ells, Ms)] == ref
    assert [_sums(nm.sample(l, M)) for l, M in zip(ells, Ms)] == ref2
    nm.cleanup_nestedmc_sampler()
    _shift_setup(nm, oracle_f64, 1, form)
    assert [_sums(r) for r in nm.sample_levels(ells, Ms)] == ref
    assert [_sums(nm.sample(l, M)) for l, M in zip(ells, Ms)] == ref2
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_mlmc_driver_sweeps_match_lone_calls(nm, oracle_f64):
    """The native MLMC driver in the batched form samples each step's levels as
    one sweep; run over lone level calls instead (the callback driver, same
    control flow) it makes the same (l, M) calls and gets the same sums."""
    _shift_setup(nm, oracle_f64, 1, nm.FORM_MIXED)
    a = nm.mlmc_run(2e-4, 2000)
    nm.cleanup_nestedmc_sampler()
    _shift_setup(nm, oracle_f64, 1, nm.FORM_MIXED)
    calls = []

    def level(ell, M):
        calls.append((ell, M))
        r = nm.sample(ell, M)
        return dict(sums_fine=r["fsums"], sums_diff=r["diffsums"], work=r["work"], M=r["M"])

    b = nm.mlmc_run_cb(level, 2e-4, 2000)
    nm.cleanup_nestedmc_sampler()
    assert len(calls) > 3
    assert len(a["levels"]) == len(b["levels"])
    for la, lb in zip(a["levels"], b["levels"]):
        assert la["M"] == lb["M"]
        assert list(la["sums_fine"]) == list(lb["sums_fine"])
        assert list(la["sums_diff"]) == list(lb["sums_diff"])
        assert la["work"] == lb["work"]
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("form", [1, 2])
def test_batched_chunked_level_calls(nm, oracle_f64, form, monkeypatch):
    """A level call larger than the chunk runs as several launch chains (here
    chunks of 97 outer samples, the last one ragged, and the continuing
    offset): per-sample results equal the unchunked oracle's, sums agree."""
    monkeypatch.setenv("NMLMC_BATCH_CHUNK", "97")
    orc = _shift_setup(nm, oracle_f64, 1, form)
    # Without an explicit offset each call continues after the level's last one.
    for ell, M, continued_from in ((0, 500, 0), (2, 250, 0), (0, 3, 500)):
        g = nm.sample(ell, M, trace=True)
        o = orc.sample(ell, M, keying=1, outer_offset=continued_from, trace=True)
        _cmp(g, o) if form == 1 else _cmp_mixed(g, o)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_batched_reference_model_nonantithetic(nm, oracle_f64):
    """Batched keying with the reference model and the plain (non-antithetic)
    fine/coarse estimator."""
    nm.init_nestedmc_sampler(1, 32, 4, 1.5, False)
    nm.set_form(nm.FORM_BATCHED)
    orc = oracle_f64.new(1, 32, 4, 1.5, False)
    ml = _ref_max_loss(orc)
    nm.model_init(1, 0.02, ml, 9)
    orc.model_init(0, 1, 0.02, ml, 9)
    for ell, M in ((0, 200), (2, 150)):
        _cmp(nm.sample(ell, M, outer_offset=0, trace=True),
             orc.sample(ell, M, keying=1, outer_offset=0, trace=True))
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def _cmp_mixed(g, o):
    """The mixed form against FP64 on the same draws: every refinement decision
    (N_f, N_c) and every indicator equal (decisions within the FP32 error of
    their threshold are rechecked in FP64, near-threshold means redone), hence
    the draw counts and work too; sums of the indicators' moments to 1e-12."""
    assert np.array_equal(g["nf"], o["nf"]), int((g["nf"] != o["nf"]).sum())
    assert np.array_equal(g["nc"], o["nc"]), int((g["nc"] != o["nc"]).sum())
    assert np.array_equal(g["phi_f"], o["phi_f"])
    assert np.array_equal(g["phi_c"], o["phi_c"])
    assert np.array_equal(g["stats"], o["stats"]), (g["stats"], o["stats"])
    np.testing.assert_allclose(g["work"], o["work"], rtol=1e-12)
    np.testing.assert_allclose(g["fsums"], o["fsums"], rtol=RTOL, atol=1e-300)
    np.testing.assert_allclose(g["diffsums"], o["diffsums"], rtol=RTOL, atol=1e-300)


@pytest.mark.parametrize("redo_all", [False, True])
def test_mixed_form_agrees_with_fp64(nm, oracle_f64, redo_all, monkeypatch):
    """Mixed FP32+FP64 form against the FP64 oracle with the same keying:
    refinement decisions, counts and indicators all equal the FP64 ones.  With
    redo_all (NMLMC_MIXED_REDO_ALL) every outer sample takes the FP64 redo: at
    level 0 twenty thousand small redos side by side, at level 5 finals of up
    to 32768 units split over 128 blocks each."""
    if redo_all:
        monkeypatch.setenv("NMLMC_MIXED_REDO_ALL", "1")
    orc = _shift_setup(nm, oracle_f64, 1, nm.FORM_MIXED)
    for ell, M in ((0, 20000), (2, 5000), (5, 500)):
        g = nm.sample(ell, M, outer_offset=0, trace=True)
        o = orc.sample(ell, M, keying=1, outer_offset=0, trace=True)
        reev, flips, rck = nm.mixed_stats()
        print("ell", ell, "re-evaluated", reev, "flips", flips, "rechecked decisions", rck, flush=True)
        _cmp_mixed(g, o)
        if redo_all:
            assert reev == M
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("sliced", ["1", "0"])
@pytest.mark.parametrize("form", [1, 2])
def test_xcd_slicing_keeps_results(nm, oracle_f64, sliced, form, monkeypatch):
    """The batched forms' passes over outer samples, test entries and final
    entries split into one slice per XCD or not (NMLMC_XCD_SLICES; the model's
    default is unsliced): per-sample results equal the oracle's batched keying
    either way."""
    monkeypatch.setenv("NMLMC_XCD_SLICES", sliced)
    orc = _shift_setup(nm, oracle_f64, 1, form)
    for ell, M in ((0, 20000), (3, 900)):
        g = nm.sample(ell, M, outer_offset=0, trace=True)
        o = orc.sample(ell, M, keying=1, outer_offset=0, trace=True)
        if form == nm.FORM_MIXED:
            _cmp_mixed(g, o)
        else:
            _cmp(g, o)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("anti", [True, False])
def test_mixed_reference_rule_and_model(nm, oracle_f64, anti):
    """Mixed form with the reference's adaptive rule (N = delta^-r tmp, r = 1.5)
    and the reference model (two normals per inner sample), antithetic and
    plain fine/coarse estimators: decisions and indicators equal FP64's."""
    nm.init_nestedmc_sampler(1, 32, 4, 1.5, anti)
    nm.set_form(nm.FORM_MIXED)
    orc = oracle_f64.new(1, 32, 4, 1.5, anti)
    ml = _ref_max_loss(orc)
    nm.model_init(0, 0.02, ml, 19)
    orc.model_init(0, 0, 0.02, ml, 19)
    for ell, M in ((0, 4000), (1, 2000), (3, 300)):
        _cmp_mixed(nm.sample(ell, M, outer_offset=0, trace=True),
                   orc.sample(ell, M, keying=1, outer_offset=0, trace=True))
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_plain_nonadaptive_shift_sweep(nm, oracle_f64):
    """The CPU-path workload on the GPU: shift model (X = Y + Z, threshold 2),
    plain FP64 form, adaptive refinement off (N = 32 2^l), antithetic, seed 1,
    8 virtual threads as with 8 OpenMP threads, levels 0..8 with
    M_l = ceil(M0 2^-1.5l) on a reduced M0: per-sample phi, counts and the
    level sums equal the oracle's (reference keying) call after call."""
    nm.init_nestedmc_sampler(8, 32, 2, 0.0, True)
    nm.model_init_ex(nm.MODEL_SHIFT, 0, 0.5, 2.0, 1)
    orc = oracle_f64.new(8, 32, 2, 0.0, True)
    orc.model_init(1, 0, 0.5, 2.0, 1)
    for ell in range(9):
        M = int(math.ceil(2e4 * 2.0 ** (-1.5 * ell)))
        g = nm.sample(ell, M, trace=True)
        o = orc.sample(ell, M, trace=True)
        _cmp(g, o)
        assert np.all(g["nf"] == 32 << ell)
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("form", [0, 1, 2])
def test_shift_model_closed_form(nm, form):
    """P(E[X|Y] > 2) = 1 - Phi(2) for the shift model: the MLMC telescoping sum
    over levels 0..5 lies within 4 standard errors (plus the level-5 bias bound)
    of the closed form."""
    nm.init_nestedmc_sampler(4096 if form == 0 else 1, 32, 2, 0.0, True)
    nm.set_adaptive_rule(1, 6)
    nm.set_form(form)
    nm.model_init_ex(nm.MODEL_SHIFT, 0, 0.5, 2.0, 123)
    exact = 0.5 * math.erfc(2 / math.sqrt(2))
    est, var = 0.0, 0.0
    for ell, M in ((0, 400000), (1, 200000), (2, 100000), (3, 50000), (4, 25000), (5, 12000)):
        o = nm.sample(ell, M)
        s = o["fsums"] if ell == 0 else o["diffsums"]
        mean = s[0] / M
        est += mean
        var += (s[1] / M - mean * mean) / M
        last = abs(mean)
    assert abs(est - exact) < 4 * math.sqrt(var) + 2 * last, (est, exact, math.sqrt(var), last)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)
