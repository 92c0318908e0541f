"""A batch of portfolio problems in one launch chain per level call: each
instance's results equal those of the same problem staged alone with the same
seed and sampled with the same level calls (per-sample draws are keyed by the
instance's seed and outer ids, so the level sums are the same numbers)."""
import importlib

import numpy as np
import pytest

from portfolio_util import synthetic, to_library, to_oracle

pytestmark = pytest.mark.gpu

SCAL = dict(r=0.05, tau=0.02, rho=0.5, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5)


def _books(nm):
    """Three Exact-only books of different sizes and dimensions, the second
    with a Cholesky factor."""
    out = []
    for k, (P, Q, chol) in enumerate(((120, 3, False), (400, 7, True), (60, 2, False))):
        d = synthetic(P=P, Q=Q, counts=(0, P, 0), digitals=True, seed=11 + k)
        sc = dict(SCAL, max_loss=0.4 * (k + 1), subsample=1, use_delta_cv=True, use_antithetic_risk=True,
                  phi_type=0, phi_size=1)
        L = None
        if chol:
            rs = np.random.RandomState(5)
            A = rs.uniform(0, 0.5, (Q, 2))
            C = A @ A.T
            np.fill_diagonal(C, 1.0)
            L = np.linalg.cholesky(C)
        out.append((d, sc, L, 100 + 7 * k))
    return out


def _alone(nm, form, book, calls):
    d, sc, L, seed = book
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(form)
    pf, _, _ = to_library(nm, d, **sc)
    nm.portfolio_init(pf, seed)
    if L is not None:
        nm.portfolio_set_correlation(L)
    res = [nm.sample(ell, M) for ell, M in calls]
    nm.cleanup_nestedmc_sampler()
    return res


def _key(r):
    return (tuple(r["fsums"]), tuple(r["diffsums"]), r["work"], tuple(int(x) for x in r["stats"]), r["M"])


@pytest.mark.parametrize("chunk", [None, "700"])
@pytest.mark.parametrize("form", [1, 2])
def test_batch_matches_instances_alone(nm, form, chunk, monkeypatch):
    """Three instances (different d, one with a Cholesky factor, different
    thresholds and seeds) in one chain per level call, over levels 0..2 and a
    repeated level (continuing outer ids); with a chunk of 700 outer samples the
    instances straddle chunk boundaries.  Level sums, work and counts equal the
    lone runs' (indicator sums and counts are integers; the work sums of
    integers too)."""
    if chunk:
        monkeypatch.setenv("NMLMC_BATCH_CHUNK", chunk)
    books = _books(nm)
    Ms = [[600, 250, 900], [200, 90, 300], [80, 40, 120], [500, 100, 50]]
    ells = [0, 1, 2, 0]
    alone = [_alone(nm, form, b, [(e, M[i]) for e, M in zip(ells, Ms)]) for i, b in enumerate(books)]
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(form)
    nm.batch_begin(len(books))
    keep = []
    for i, (d, sc, L, seed) in enumerate(books):
        pf, _, _ = to_library(nm, d, **sc)
        keep.append(pf)
        nm.batch_stage(i, pf, seed, L)
    for c, (ell, M) in enumerate(zip(ells, Ms)):
        got = nm.batch_sample(ell, M)
        for i in range(len(books)):
            assert _key(got[i]) == _key(alone[i][c]), (c, i)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("form", [1, 2])
def test_batch_autotune_keeps_results(nm, form):
    """The batch launch-shape autotuner times its candidates on outer ids set
    apart for tuning: the batch calls after it equal the instances' lone calls
    (the instances' outer ids did not move), whatever shape it chose."""
    books = _books(nm)
    Ms = [[600, 250, 900], [200, 90, 300]]
    ells = [0, 2]
    alone = [_alone(nm, form, b, [(e, M[i]) for e, M in zip(ells, Ms)]) for i, b in enumerate(books)]
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(form)
    nm.batch_begin(len(books))
    keep = []
    for i, (d, sc, L, seed) in enumerate(books):
        pf, _, _ = to_library(nm, d, **sc)
        keep.append(pf)
        nm.batch_stage(i, pf, seed, L)
    for ell, M in zip(ells, Ms):
        S, B, ms = nm.batch_autotune_level(ell, M)
        assert S in (4, 8, 16, 32, 64, 128) and B in (4, 8, 16) and ms > 0
    for c, (ell, M) in enumerate(zip(ells, Ms)):
        got = nm.batch_sample(ell, M)
        for i in range(len(books)):
            assert _key(got[i]) == _key(alone[i][c]), (c, i)
    with pytest.raises(nm.NmlmcError):
        nm.batch_autotune_level(0, [100, 100])  # two entries for three instances
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def _asian_book(nm):
    """A general book: Exact options and Unbiased ones of which about half are
    arithmetic-average Asian calls / puts."""
    d = synthetic(P=36, Q=3, counts=(0, 20, 16), digitals=True, seed=31)
    rs = np.random.RandomState(32)
    t = d["type"].copy()
    for j in range(20, 36):
        if rs.rand() < 0.5:
            t[j] = 3 if rs.rand() < 0.5 else -3
    d["type"] = t
    sc = dict(SCAL, max_loss=0.01, subsample=1, use_delta_cv=True, use_antithetic_risk=True,
              phi_type=0, phi_size=1)
    return d, sc, None, 17


@pytest.mark.parametrize("form", [1, 2])
def test_batch_with_asian_book_matches_alone(nm, form):
    """A book with Unbiased options (Asians among them) between two Exact-only
    books: every instance's level sums, work and counts equal its lone run's."""
    b = _books(nm)
    books = [b[0], _asian_book(nm), b[2]]
    Ms = [[300, 60, 200], [100, 40, 50], [40, 20, 30]]
    ells = [0, 1, 2]
    alone = [_alone(nm, form, bk, [(e, M[i]) for e, M in zip(ells, Ms)]) for i, bk in enumerate(books)]
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(form)
    nm.batch_begin(len(books))
    keep = []
    for i, (d, sc, L, seed) in enumerate(books):
        pf, _, _ = to_library(nm, d, **sc)
        keep.append(pf)
        nm.batch_stage(i, pf, seed, L)
    for c, (ell, M) in enumerate(zip(ells, Ms)):
        got = nm.batch_sample(ell, M)
        for i in range(len(books)):
            assert _key(got[i]) == _key(alone[i][c]), (c, i)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


@pytest.mark.parametrize("form", [1, 2])
def test_batch_books_match_oracle(nm, oracle_f64, form):
    """Six books of the batched-risk workload (d in {10, 50, 100}, P
    log-uniform in [10^3, 10^5], 20% Asians by the Unbiased method, 20%
    digitals, factor correlation, pilot-quantile thresholds) sampled together,
    each instance's level sums, work and counts against the oracle run on that
    book alone with the batched keying (its seed, continuing outer ids): the
    batch is checked against the reference algorithm itself, not only against
    the library's own lone runs."""
    cfg = importlib.import_module(nm.__name__.rsplit(".", 1)[0] + ".configs")
    books = cfg.batch_books(6, seed=7, n_pilot=100, asian_frac=0.2)
    calls = [(0, [160, 90, 200, 120, 60, 100]), (1, [80, 50, 60, 40, 70, 30]),
             (2, [24, 30, 16, 20, 12, 28]), (1, [20, 10, 30, 10, 20, 10])]
    nm.init_nestedmc_sampler(1, 64, 2, 0.0, True)
    nm.set_adaptive_rule(1, 4)
    nm.set_form(form)
    nm.batch_begin(len(books))
    keep, orcs = [], []
    for i, (a, L, sc, q) in enumerate(books):
        pf, start, delta = to_library(nm, a, **sc)
        keep.append(pf)
        nm.batch_stage(i, pf, 300 + i, L)
        o = oracle_f64.__class__("f64").new(1, 64, 2, 0.0, True, adapt_rule=1, cap_log2=4)
        o.portfolio_init(to_oracle(a, start, delta, **sc), 300 + i, L)
        orcs.append(o)
    offsets = {}
    for ell, Ms in calls:
        print("batch level", ell, Ms, flush=True)
        got = nm.batch_sample(ell, Ms)
        for i, o in enumerate(orcs):
            off = offsets.get((i, ell), 0)
            r = o.sample(ell, Ms[i], keying=1, outer_offset=off)
            offsets[(i, ell)] = off + Ms[i]
            g = got[i]
            assert np.array_equal(g["stats"], r["stats"]), (i, ell, g["stats"], r["stats"])
            np.testing.assert_allclose(g["work"], r["work"], rtol=1e-12)
            np.testing.assert_allclose(g["fsums"], r["fsums"], rtol=1e-9, atol=1e-300)
            np.testing.assert_allclose(g["diffsums"], r["diffsums"], rtol=1e-9, atol=1e-300)
        if form == 2:
            print("mixed stats (redone, flips, rechecked)", nm.mixed_stats(), flush=True)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)


def test_batch_errors(nm):
    """A batch needs staged instances and a batched form; at most 64 instances."""
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_form(nm.FORM_MIXED)
    with pytest.raises(nm.NmlmcError):
        nm.batch_begin(65)
    nm.batch_begin(2)
    d, sc, L, seed = _books(nm)[0]
    pf, _, _ = to_library(nm, d, **sc)
    nm.batch_stage(0, pf, seed)
    with pytest.raises(nm.NmlmcError):
        nm.batch_sample(0, [10, 10])  # instance 1 not staged
    nm.batch_stage(1, pf, seed + 1)
    with pytest.raises(nm.NmlmcError):
        nm.batch_sample(0, [10])  # one entry for a batch of two
    with pytest.raises(nm.NmlmcError):
        nm.batch_mlmc_run([1e-2, 1e-2, 1e-2], 100)  # three tolerances for a batch of two
    nm.set_form(nm.FORM_PLAIN)
    with pytest.raises(nm.NmlmcError):
        nm.batch_sample(0, [10, 10])
    nm.cleanup_nestedmc_sampler()


def test_batch_mlmc_drivers_match_runs_alone(nm):
    """The MLMC drivers of a batch run together give every instance the run it
    gets alone (same levels, sample counts and sums), each to its own
    tolerance; the estimates are finite probabilities."""
    books = _books(nm)
    tols = [4e-3, 6e-3, 5e-3]
    alone = []
    for (d, sc, L, seed), tol in zip(books, tols):
        nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
        nm.set_adaptive_rule(1, 3)
        nm.set_form(nm.FORM_MIXED)
        pf, _, _ = to_library(nm, d, **sc)
        nm.portfolio_init(pf, seed)
        if L is not None:
            nm.portfolio_set_correlation(L)
        alone.append(nm.mlmc_run(tol, 200))
        nm.cleanup_nestedmc_sampler()
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(nm.FORM_MIXED)
    nm.batch_begin(len(books))
    keep = []
    for i, (d, sc, L, seed) in enumerate(books):
        pf, _, _ = to_library(nm, d, **sc)
        keep.append(pf)
        nm.batch_stage(i, pf, seed, L)
    got = nm.batch_mlmc_run(tols, 200)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)
    for a, g in zip(alone, got):
        assert len(a["levels"]) == len(g["levels"]) and a["starting_level"] == g["starting_level"]
        for la, lg in zip(a["levels"], g["levels"]):
            assert la["M"] == lg["M"]
            assert la["sums_fine"] == lg["sums_fine"] and la["sums_diff"] == lg["sums_diff"]
            assert la["work"] == lg["work"]
        e = nm.mlmc_estimate(g)
        assert 0.0 <= e <= 1.0


@pytest.mark.parametrize("form", [1, 2])
def test_portfolio_level_sweep_and_autotune(nm, form):
    """Portfolio problems in a level sweep (one captured graph) and after the
    launch-shape autotuner: the same level sums as lone level calls on the same
    outer samples (the correlated book goes through the FP64 matrix-core outer
    draw)."""
    d, sc, L, seed = _books(nm)[1]
    calls = [(0, 700), (1, 300), (2, 120)]
    ref = _alone(nm, form, (d, sc, L, seed), calls)
    nm.init_nestedmc_sampler(1, 16, 2, 0.0, True)
    nm.set_adaptive_rule(1, 3)
    nm.set_form(form)
    pf, _, _ = to_library(nm, d, **sc)
    nm.portfolio_init(pf, seed)
    nm.portfolio_set_correlation(L)
    for ell, M in calls:
        nm.autotune_level(ell, M)
    got = nm.sample_levels([e for e, _ in calls], [M for _, M in calls])
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)
    for g, r in zip(got, ref):
        assert _key(g) == _key(r)


@pytest.mark.parametrize("form", [1, 2])
def test_batch_level_coupled_books_match_oracle(nm, oracle_f64, form):
    """The batched-risk workload as specified: books with 20% arithmetic Asians
    on level-coupled Milstein paths, sampled together; each instance's level
    sums, work and counts against the oracle on that book alone."""
    cfg = importlib.import_module(nm.__name__.rsplit(".", 1)[0] + ".configs")
    books = cfg.batch_books(4, seed=9, n_pilot=100, asian_frac=0.2)
    calls = [(0, [120, 90, 60, 100]), (1, [60, 50, 40, 30]), (2, [20, 24, 16, 12]), (3, [10, 12, 8, 6])]
    nm.init_nestedmc_sampler(1, 64, 2, 0.0, True)
    nm.set_adaptive_rule(1, 4)
    nm.set_form(form)
    nm.set_level_coupled_paths(True)
    nm.batch_begin(len(books))
    keep, orcs = [], []
    for i, (a, L, sc, q) in enumerate(books):
        pf, start, delta = to_library(nm, a, **sc)
        keep.append(pf)
        nm.batch_stage(i, pf, 500 + i, L)
        o = oracle_f64.__class__("f64").new(1, 64, 2, 0.0, True, adapt_rule=1, cap_log2=4)
        o.portfolio_init(to_oracle(a, start, delta, **sc), 500 + i, L)
        o.set_coupled(True)
        orcs.append(o)
    offsets = {}
    for ell, Ms in calls:
        print("batch level", ell, Ms, flush=True)
        got = nm.batch_sample(ell, Ms)
        for i, o in enumerate(orcs):
            off = offsets.get((i, ell), 0)
            r = o.sample(ell, Ms[i], keying=1, outer_offset=off)
            offsets[(i, ell)] = off + Ms[i]
            g = got[i]
            assert np.array_equal(g["stats"], r["stats"]), (i, ell, g["stats"], r["stats"])
            np.testing.assert_allclose(g["work"], r["work"], rtol=1e-12)
            np.testing.assert_allclose(g["fsums"], r["fsums"], rtol=1e-9, atol=1e-300)
            np.testing.assert_allclose(g["diffsums"], r["diffsums"], rtol=1e-9, atol=1e-300)
    nm.set_level_coupled_paths(False)
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)
