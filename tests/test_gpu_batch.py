"""A batch of portfolio problems in one launch chain per level call: each
instance's results equal those of the same problem staged alone with the same
seed and sampled with the same level calls (per-sample draws are keyed by the
instance's seed and outer ids, so the level sums are the same numbers)."""
import importlib

import numpy as np
import pytest

from portfolio_util import synthetic, to_library

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
