"""GPU parity of the portfolio problem: the HIP library against the oracle on
the same portfolios and seeds.

Plain FP64 form (reference keying): per-sample phi_f, phi_c, N_f, N_c and the
level sums over consecutive calls, on the reference's own generated portfolio
(BS, Exact and Unbiased/Milstein options) and on synthetic ones (digitals,
Cholesky correlation), for every sub-sampling mode and phi type.  Batched FP64
form: the same against the oracle's batched keying.  Mixed form: indicators
equal to FP64 wherever the refinement decisions agree.

Tolerances: counts, N_f, N_c and inner_stats exact.  FP64 sums and phi values:
1e-12 relative for BS/Exact portfolios (device libm and summation order), with
an absolute floor of 1e-12 times the largest sum.  Portfolios holding Unbiased
(Milstein) options get 1e-9 relative: their estimator divides a fine-minus-
coarse path difference by the sampled level's probability, so a 1-ulp
difference between the device's and glibc's exp/log/sqrt is amplified by that
cancellation and by 1/p (observed: up to 4e-11)."""
import numpy as np
import pytest

from portfolio_util import reference_generated, synthetic, to_library, to_oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12
RTOL_UNBIASED = 1e-9


def _cmp(g, o, trace=True, rtol=RTOL, exact_work=False):
    assert np.array_equal(g["stats"], o["stats"]), (g["stats"], o["stats"])
    if exact_work:
        assert g["work"] == o["work"]
    else:
        np.testing.assert_allclose(g["work"], o["work"], rtol=1e-12)
    scale = max(1.0, float(np.max(np.abs(o["fsums"]))))
    atol = rtol * scale
    np.testing.assert_allclose(g["fsums"], o["fsums"], rtol=rtol, atol=atol)
    np.testing.assert_allclose(g["diffsums"], o["diffsums"], rtol=rtol, atol=atol)
    if trace:
        assert np.array_equal(g["nf"], o["nf"])
        assert np.array_equal(g["nc"], o["nc"])
        np.testing.assert_allclose(g["phi_f"], o["phi_f"], rtol=rtol, atol=0.1 * rtol)
        np.testing.assert_allclose(g["phi_c"], o["phi_c"], rtol=rtol, atol=0.1 * rtol)


def _cmp_mixed(g, o, rtol=RTOL):
    """The mixed form against FP64 on the same draws: refinement decisions
    (N_f, N_c) and indicators equal (uncertain decisions are rechecked in FP64,
    near-threshold means redone), hence the draw counts; work and the sums of
    the indicators' moments to rtol."""
    assert np.array_equal(g["nf"], o["nf"]), int((g["nf"] != o["nf"]).sum())
    assert np.array_equal(g["nc"], o["nc"]), int((g["nc"] != o["nc"]).sum())
    assert np.array_equal(g["phi_f"], o["phi_f"]), int((g["phi_f"] != o["phi_f"]).sum())
    assert np.array_equal(g["phi_c"], o["phi_c"]), int((g["phi_c"] != o["phi_c"]).sum())
    assert np.array_equal(g["stats"], o["stats"]), (g["stats"], o["stats"])
    np.testing.assert_allclose(g["work"], o["work"], rtol=rtol)
    np.testing.assert_allclose(g["fsums"], o["fsums"], rtol=rtol, atol=1e-300)
    np.testing.assert_allclose(g["diffsums"], o["diffsums"], rtol=rtol, atol=1e-300)


SCAL = dict(r=0.05, rho=0.2, tau=0.02, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5,
            use_coarse_delta_cv=False, use_milstein=True)


def _setup(nm, orc, d, cores, N0, beta, r_adapt, scal, seed, form=0, chol=None, rule=None):
    nm.init_nestedmc_sampler(cores, N0, beta, r_adapt, True)
    if rule:
        nm.set_adaptive_rule(*rule)
    nm.set_form(form)
    pf, start, delta = to_library(nm, d, **scal)
    nm.portfolio_init(pf, seed)
    if chol is not None:
        nm.portfolio_set_correlation(chol)
    o = orc.new(cores, N0, beta, r_adapt, True,
                **({"adapt_rule": rule[0], "cap_log2": rule[1]} if rule else {}))
    o.portfolio_init(to_oracle(d, start, delta, **scal), seed, chol)
    return pf, o


@pytest.mark.parametrize("sub,cv,phi,ps,cores", [
    (1, True, 0, 1, 1), (1, True, 1, 1, 4), (-1, False, 1, 1, 2), (0, True, 0, 1, 1),
    (1, True, 2, 2, 3), (1, False, 3, 3, 1)])
def test_plain_reference_portfolio(nm, oracle_f64, sub, cv, phi, ps, cores):
    """The reference's generated portfolio (BS / Exact / Unbiased-Milstein)."""
    print("case", sub, cv, phi, ps, cores, flush=True)
    d = reference_generated()
    scal = dict(SCAL, max_loss=0.02, subsample=sub, use_delta_cv=cv, use_antithetic_risk=cv,
                phi_type=phi, phi_size=ps)
    pf, o = _setup(nm, oracle_f64, d, cores, 8, 4, 1.5, scal, 3)
    for ell, M in ((0, 9), (1, 6), (2, 3)):
        g = nm.sample(ell, M, trace=True, phi_size=pf.phi_size)
        _cmp(g, o.sample(ell, M, trace=True), rtol=RTOL_UNBIASED)
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("counts,digitals,cores", [((10, 30, 0), True, 1), ((0, 40, 0), False, 8),
                                                   ((20, 20, 0), True, 64)])
def test_plain_fast_path_portfolio(nm, oracle_f64, counts, digitals, cores):
    """Sub-sampled BS + Exact portfolios: the wave-parallel inner loop."""
    d = synthetic(counts=counts, digitals=digitals, seed=len(counts) + cores)
    scal = dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=1, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, cores, 32, 4, 1.5, scal, 5)
    for ell, M in ((0, 200), (1, 120), (3, 30), (2, 50)):
        _cmp(nm.sample(ell, M, trace=True), o.sample(ell, M, trace=True))
    nm.cleanup_nestedmc_sampler()


def test_plain_cholesky_and_drop_in_entry(nm, oracle_f64):
    """General correlation (Cholesky factor) and the portfolio_sample entry."""
    d = synthetic(Q=5, counts=(5, 35, 0), seed=4)
    C = 0.3 * np.ones((5, 5)) + 0.7 * np.eye(5)
    L = np.linalg.cholesky(C)
    scal = dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, 4, 16, 4, 1.5, scal, 7, chol=L)
    for ell, M in ((0, 50), (2, 20)):
        g = nm.portfolio_sample(ell, M)
        _cmp(g, o.sample(ell, M), trace=False)
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("form,which", [(1, "reference"), (1, "synthetic"), (1, "cholesky"),
                                        (2, "reference"), (2, "synthetic"), (2, "cholesky")])
def test_batched_fp64_matches_oracle(nm, oracle_f64, form, which):
    """Batched FP64 form (phi = Max) and mixed form (phi = Indicator) against
    the oracle's batched keying, under the reference's adaptive rule, on the
    reference's generated book (BS, Exact, Unbiased) and on a BS + Exact book
    with digitals (whose mixed form takes the FP32 Exact records of the general
    kernel), with and without a Cholesky factor."""
    if which == "reference":
        d = reference_generated()
        chol = None
    else:
        d = synthetic(Q=6, counts=(10, 50, 0), seed=11)
        chol = np.linalg.cholesky(0.25 * np.ones((6, 6)) + 0.75 * np.eye(6)) if which == "cholesky" else None
    scal = dict(SCAL, max_loss=0.02, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=1 if form == 1 else 0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, 1, 8, 4, 1.5, scal, 9, form=form, chol=chol)
    rtol = RTOL_UNBIASED if which == "reference" else RTOL
    for ell, M, off in ((0, 300, 0), (1, 200, 0), (3, 40, 0), (1, 100, 200)):
        g = nm.sample(ell, M, outer_offset=off, trace=True)
        r = o.sample(ell, M, keying=1, outer_offset=off, trace=True)
        _cmp(g, r, rtol=rtol) if form == 1 else _cmp_mixed(g, r, rtol=rtol)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def _with_asians(seed):
    """Exact options plus Unbiased ones of which about half are arithmetic-
    average Asian calls / puts (types 3 / -3)."""
    d = synthetic(Q=3, counts=(0, 20, 16), seed=seed)
    n1 = int(d["counts"][1])
    rs = np.random.RandomState(seed + 1)
    t = d["type"].copy()
    for j in range(n1, len(t)):
        if rs.rand() < 0.5:
            t[j] = 3 if rs.rand() < 0.5 else -3
    d["type"] = t
    return d


@pytest.mark.parametrize("form,cores,cv", [(0, 1, True), (0, 3, False), (1, 1, True), (2, 1, False)])
def test_asian_options_match_oracle(nm, oracle_f64, form, cores, cv):
    """Arithmetic-average Asian options (the Unbiased method's Milstein paths
    with their averages over (tau, T]) beside Exact and European Unbiased
    options: plain form against the reference keying, batched and mixed forms
    against the batched keying; per-sample results (the mixed form redoes a
    near-threshold sample in FP64, and the Unbiased terms are FP64 anyway)."""
    d = _with_asians(31)
    assert np.any(np.abs(d["type"]) == 3)
    scal = dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=cv, use_antithetic_risk=True,
                phi_type=1 if form < 2 else 0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, cores, 8, 4, 1.5 if form == 0 else 0.0, scal, 17, form=form,
                   rule=None if form == 0 else (1, 3))
    for ell, M in ((0, 40), (1, 30), (2, 12)):
        if form:
            g = nm.sample(ell, M, outer_offset=0, trace=True)
            r = o.sample(ell, M, keying=1, outer_offset=0, trace=True)
        else:
            g = nm.sample(ell, M, trace=True)
            r = o.sample(ell, M, trace=True)
        if form == 2:
            _cmp_mixed(g, r, rtol=RTOL_UNBIASED)
        else:
            _cmp(g, r, rtol=RTOL_UNBIASED)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("form", [1, 2])
def test_deep_unbiased_levels_match_oracle(nm, oracle_f64, form):
    """Unbiased samples with level >= 2 are computed by a whole wave, each path
    split into chunks of coarse steps whose factors are chained in order.  With
    the level law P(L >= l) = 4^(-1.1 l) about one Unbiased sample in twenty is
    that deep (paths of 128 to 10^5 steps).  The chained product rounds
    differently from the oracle's one-lane recurrence, and a deep sample's
    weight 4^(1.1 L) / p(option) magnifies that rounding: per-sample results and
    level sums agree to 1e-7 relative (the mixed form's indicators exactly)."""
    d = synthetic(Q=3, counts=(0, 10, 30), seed=41)
    rs = np.random.RandomState(42)
    t = d["type"].copy()
    for j in range(10, 40):
        if rs.rand() < 0.5:
            t[j] = 3 if rs.rand() < 0.5 else -3
    d["type"] = t
    scal = dict(SCAL, dt_prob_expo=-1.1, max_loss=0.01, subsample=1, use_delta_cv=True,
                use_antithetic_risk=True, phi_type=1 if form == 1 else 0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, 1, 8, 4, 0.0, scal, 23, form=form, rule=(1, 3))
    for ell, M in ((0, 60), (1, 40), (2, 16)):
        g = nm.sample(ell, M, outer_offset=0, trace=True)
        r = o.sample(ell, M, keying=1, outer_offset=0, trace=True)
        if form == 2:
            _cmp_mixed(g, r, rtol=1e-7)
        else:
            _cmp(g, r, rtol=1e-7)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_asian_options_need_the_unbiased_method(nm):
    """Asian options have no exact terminal law and no Black-Scholes value:
    staged as Exact or BS options they are refused."""
    d = synthetic(Q=2, counts=(0, 6, 0), seed=3)
    d["type"] = d["type"].copy()
    d["type"][0] = 3
    nm.init_nestedmc_sampler(1, 8, 2, 0.0, True)
    pf, _, _ = to_library(nm, d, **dict(SCAL, max_loss=0.0, subsample=1, phi_type=0, phi_size=1))
    with pytest.raises(nm.NmlmcError):
        nm.portfolio_init(pf, 1)
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("redo_all", [False, True])
def test_mixed_portfolio_agrees_with_fp64(nm, oracle_f64, redo_all, monkeypatch):
    """Mixed FP32+FP64 form on an Exact portfolio with digitals: refinement
    decisions and indicators equal the FP64 ones.  With redo_all
    (NMLMC_MIXED_REDO_ALL) every outer sample takes the FP64 redo, up to finals
    of 16384 inner samples split over 64 blocks at level 5."""
    if redo_all:
        monkeypatch.setenv("NMLMC_MIXED_REDO_ALL", "1")
    d = synthetic(Q=4, counts=(0, 200, 0), seed=21)
    scal = dict(SCAL, max_loss=0.005, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, 1, 32, 2, 0.0, scal, 13, form=2, rule=(1, 4))
    calls = ((0, 3000), (1, 1500), (3, 300)) + (((5, 40),) if redo_all else ())
    for ell, M in calls:
        g = nm.sample(ell, M, outer_offset=0, trace=True)
        r = o.sample(ell, M, keying=1, outer_offset=0, trace=True)
        reev, flips, rck = nm.mixed_stats()
        print("ell", ell, "mixed stats (redone, flips, rechecked)", (reev, flips, rck), flush=True)
        _cmp_mixed(g, r)
        if redo_all:
            assert reev == M
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_portfolio_errors(nm):
    """Host-side validation: unsorted probabilities are rejected with a status."""
    nm.init_nestedmc_sampler(1, 8, 4, 0.0, True)
    d = synthetic(counts=(0, 5, 0), seed=1)
    d["prob"] = d["prob"][::-1].copy()
    pf, _, _ = to_library(nm, d, **dict(SCAL, max_loss=0.0))
    with pytest.raises(nm.NmlmcError):
        nm.portfolio_init(pf, 1)
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("cores", [1, 3, 64])
def test_delta_unbiased_matches_oracle(nm, oracle_f64, cores):
    """portfolio_sample_delta_unbiased on the reference's generated portfolio:
    consecutive calls from the persistent inner streams, then a level call on
    the same streams.  Work exact; moments to the Unbiased tolerance."""
    d = reference_generated()
    scal = dict(SCAL, max_loss=0.02, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, cores, 8, 4, 1.5, scal, 5)
    for idx, M in ((0, 40), (3, 25), (0, 7), (11, 300)):
        g_sums, g_work = nm.portfolio_sample_delta_unbiased(idx, M, 2)
        o_sums, o_work = o.sample_delta_unbiased(idx, M, 2)
        assert g_work == o_work
        np.testing.assert_allclose(g_sums, o_sums, rtol=RTOL_UNBIASED, atol=1e-12)
    _cmp(nm.sample(1, 5, trace=True), o.sample(1, 5, trace=True), rtol=RTOL_UNBIASED)
    with pytest.raises(nm.NmlmcError):
        nm.portfolio_sample_delta_unbiased(12, 10)  # the portfolio has 12 Unbiased options
    nm.cleanup_nestedmc_sampler()


def test_estimate_delta_for_unbiased(nm, oracle_f64):
    """portfolio_estimate_delta_for_unbiased: batches of 1000 delta samples per
    Unbiased option until 3 sqrt(V/M) < TOL, the mean added to delta_est;
    restated on the oracle's delta sampler."""
    d = reference_generated()
    scal = dict(SCAL, max_loss=0.02, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=0, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, 4, 8, 4, 1.5, scal, 17)
    n_unb = int(d["counts"][2])
    before = pf.column("delta_est")[-n_unb:].copy()
    TOL = 2e-3
    nm.portfolio_estimate_delta_for_unbiased(pf, TOL)
    after = pf.column("delta_est")[-n_unb:]
    expect = []
    for j in range(n_unb):
        tot, tm = np.zeros(2), 0
        while True:
            s, _ = o.sample_delta_unbiased(j, 1000, 2)
            tot += s
            tm += 1000
            V = tot[1] / tm - (tot[0] / tm) ** 2
            if 3 * np.sqrt(V / tm) < TOL:
                break
        expect.append(tot[0] / tm)
    np.testing.assert_allclose(after - before, expect, rtol=RTOL_UNBIASED, atol=1e-12)
    nm.cleanup_nestedmc_sampler()


def _configs(nm):
    import importlib
    return importlib.import_module(nm.__name__.rsplit(".", 1)[0] + ".configs")


@pytest.mark.parametrize("form,which", [(0, "small"), (1, "small"), (2, "small"), (0, "large"), (1, "large"),
                                        (2, "large")])
def test_config_shaped_portfolios(nm, oracle_f64, form, which):
    """Reduced copies of the benchmark books (10 assets with pairwise
    correlation 0.3, or 100 assets with a 5-factor correlation; Exact calls,
    puts and digitals; t3 weights for the large one) under the capped 3-sigma
    rule: plain form against the reference keying, batched and mixed against
    the batched keying; per-sample results exact (the mixed form's decisions
    and indicators equal FP64's)."""
    cfg = _configs(nm)
    if which == "small":
        a, L, sc = cfg.small_portfolio(P=2000, n_pilot=300)
        calls = ((0, 400), (1, 200), (3, 60))
    else:
        a, L, sc = cfg.large_portfolio(P=5000, n_pilot=200)
        calls = ((0, 200), (2, 60))
    pf, o = _setup(nm, oracle_f64, a, 1 if form else 64, 64, 2, 0.0, sc, 23, form=form, chol=L,
                   rule=(1, 4))
    for ell, M in calls:
        if form:
            g = nm.sample(ell, M, outer_offset=0, trace=True)
            r = o.sample(ell, M, keying=1, outer_offset=0, trace=True)
        else:
            g = nm.sample(ell, M, trace=True)
            r = o.sample(ell, M, trace=True)
        if form == 2:
            print("ell", ell, "mixed stats (redone, flips, rechecked)", nm.mixed_stats(), flush=True)
            _cmp_mixed(g, r)
        else:
            _cmp(g, r)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_plain_waves_per_virtual_thread(nm, oracle_f64, waves, monkeypatch):
    """Plain form on a book whose options are all Exact, each virtual thread on
    a block of 1, 2, 4 or 8 waves (NMLMC_PLAIN_WAVES): per-sample results equal
    the reference keying's (the samples' draws sit at known stream positions,
    so the block's lanes split them)."""
    monkeypatch.setenv("NMLMC_PLAIN_WAVES", str(waves))
    cfg = _configs(nm)
    a, L, sc = cfg.small_portfolio(P=2000, n_pilot=300)
    pf, o = _setup(nm, oracle_f64, a, 16, 64, 2, 0.0, sc, 37, form=0, chol=L, rule=(1, 4))
    for ell, M in ((0, 100), (2, 24), (4, 7)):
        _cmp(nm.sample(ell, M, trace=True), o.sample(ell, M, trace=True))
    nm.cleanup_nestedmc_sampler()


@pytest.mark.parametrize("form", [0, 1])
def test_option_search_with_ties_and_tiny_probabilities(nm, oracle_f64, form):
    """Sub-sampling probabilities spanning 16 decades, so that many exclusive-
    prefix CDF entries are equal in FP64 (ties), plus a block of equal
    probabilities: the option drawn for every inner sample matches the
    reference search (per-sample results exact)."""
    d = synthetic(Q=4, counts=(0, 300, 0), digitals=True, seed=31)
    rs = np.random.RandomState(5)
    p = np.concatenate([10.0 ** rs.uniform(-16, 0, 200), np.full(100, 1e-3)])
    order = np.argsort(-p, kind="stable")
    d["prob"] = p[order]
    scal = dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=1, phi_size=1)
    pf, o = _setup(nm, oracle_f64, d, 1 if form else 16, 16, 2, 0.0, scal, 41, form=form,
                   rule=(1, 3))
    for ell, M in ((0, 300), (2, 80)):
        if form:
            _cmp(nm.sample(ell, M, outer_offset=0, trace=True),
                 o.sample(ell, M, keying=1, outer_offset=0, trace=True))
        else:
            _cmp(nm.sample(ell, M, trace=True), o.sample(ell, M, trace=True))
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("form,segment,redo_all", [(1, 0, False), (2, 0, False), (1, 3, False), (2, 3, False),
                                                  (2, 0, True)])
def test_level_coupled_paths_match_oracle(nm, oracle_f64, form, segment, redo_all, monkeypatch):
    """Level-coupled Milstein paths for the method-2 options (Asians among them)
    beside Exact options: at level l an Asian inner sample runs
    8 2^floor(l/2) steps, and an antithetic final's coarse estimator takes the
    values on the level below's steps from the same increments.  Batched FP64
    and mixed forms against the oracle's batched keying over levels 0..4
    (with segment = 3, the final draw groups split over several lanes and
    their fine and coarse sums go through the cross-lane hand-off)."""
    if segment:
        monkeypatch.setenv("NMLMC_SEGMENT", str(segment))
    if redo_all:  # every outer sample through the sequential FP64 redo (two-valued groups)
        monkeypatch.setenv("NMLMC_MIXED_REDO_ALL", "1")
    d = _with_asians(51)
    scal = dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                phi_type=1 if form == 1 else 0, phi_size=1)
    nm.init_nestedmc_sampler(1, 8, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(form)
    nm.set_level_coupled_paths(True)
    pf, start, delta = to_library(nm, d, **scal)
    nm.portfolio_init(pf, 29)
    o = oracle_f64.new(1, 8, 2, 0.0, True, adapt_rule=1, cap_log2=3)
    o.portfolio_init(to_oracle(d, start, delta, **scal), 29)
    o.set_coupled(True)
    for ell, M in ((0, 40), (1, 30), (2, 24), (3, 16), (4, 12)):
        g = nm.sample(ell, M, outer_offset=0, trace=True)
        r = o.sample(ell, M, keying=1, outer_offset=0, trace=True)
        print("level", ell, flush=True)
        _cmp_mixed(g, r, rtol=RTOL_UNBIASED) if form == 2 else _cmp(g, r, rtol=RTOL_UNBIASED)
        if redo_all:
            assert nm.mixed_stats()[0] == M
    nm.set_level_coupled_paths(False)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_level_coupled_paths_need_batched_form(nm):
    """The plain form (the reference keying) has no level-coupled paths, and
    they need sub-sampling mode 1: both refused with a status."""
    d = _with_asians(52)
    nm.init_nestedmc_sampler(1, 8, 2, 0.0, True)
    nm.set_level_coupled_paths(True)
    pf, _, _ = to_library(nm, d, **dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=True,
                                        use_antithetic_risk=True, phi_type=0, phi_size=1))
    nm.portfolio_init(pf, 3)
    with pytest.raises(nm.NmlmcError):
        nm.sample(0, 10)
    pf0, _, _ = to_library(nm, d, **dict(SCAL, max_loss=0.01, subsample=0, use_delta_cv=True,
                                         use_antithetic_risk=True, phi_type=0, phi_size=1))
    with pytest.raises(nm.NmlmcError):
        nm.portfolio_init(pf0, 3)
    nm.set_level_coupled_paths(False)
    nm.cleanup_nestedmc_sampler()
