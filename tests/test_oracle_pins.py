"""Pins of the oracle's extension functions (those with no counterpart in the
reference, so no reference golden covers them): the Cholesky outer draw L z,
the Exact-GBM digital payoff and the capped 3-sigma refinement decision.  Each
is checked against an independent restatement (tests/rng_np.py) or a closed
form, chosen so that a plausible slip fails: a transposed factor, a wrong
normal order, a wrong digital side or drift, a wrong constant in the rule, or
a missing forced step to the cap."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import Oracle

import rng_np


def test_numpy_threefry_matches_published_kat():
    """The numpy restatement against Random123's published vectors."""
    kat = [
        ((0, 0), (0, 0), (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)),
        ((2**64 - 1, 2**64 - 1), (2**64 - 1, 2**64 - 1), (0xE02CB7C4D95D277A, 0xD06633D0893B8B68)),
        ((0x243F6A8885A308D3, 0x13198A2E03707344), (0xA4093822299F31D0, 0x082EFA98EC4E6C89),
         (0x263C7D30BB0F0AF1, 0x56BE8361D3311526)),
    ]
    for (c0, c1), (k0, k1), out in kat:
        a, b = rng_np.threefry2x64_20(c0, c1, k0, k1)
        assert (int(a), int(b)) == out


def _ncdf(x):
    return 0.5 * special.erfc(-np.asarray(x) / math.sqrt(2.0))


def _book(Q, counts, types, strikes, mats, w, stock_ind, tau, r=0.05, **kw):
    rs = np.random.RandomState(Q)
    d = dict(S0=rs.uniform(90, 110, Q), mu=rs.uniform(0.02, 0.12, Q), sigma=rs.uniform(0.15, 0.4, Q),
             counts=np.array(counts, np.uint32), strike=np.array(strikes, float),
             maturity=np.array(mats, float), start=np.zeros(len(w)), w=np.array(w, float),
             delta_est=np.zeros(len(w)), stock_ind=np.array(stock_ind, np.uint32),
             type=np.array(types, np.int32), prob=np.abs(np.array(w, float)), W_est=np.ones(len(w)),
             r=r, rho=0.0, max_loss=0.0, tau=tau, phi_type=2, phi_size=1, subsample=0,
             use_delta_cv=False, use_coarse_delta_cv=False, use_antithetic_risk=True, use_milstein=True)
    d.update(kw)
    return d


def _chol(Q):
    """Lower Cholesky factor of an equicorrelated (Q = 2: 0.6) or random SPD
    correlation matrix."""
    if Q == 2:
        return np.array([[1.0, 0.0], [0.6, 0.8]])
    rs = np.random.RandomState(7)
    B = rs.uniform(-1, 1, (Q, Q))
    C = B @ B.T + Q * np.eye(Q)
    dg = np.sqrt(np.diag(C))
    return np.linalg.cholesky(C / np.outer(dg, dg))


@pytest.mark.parametrize("Q", [2, 3])
def test_cholesky_outer_draw_against_numpy_restatement(Q):
    """R_tau of outer sample g at level l: z = the first Q normals of the
    stream keyed {seed, l << 46 | g}, W = sqrt(tau) L z (L lower triangular,
    row i uses z_0..z_i), R_i = S0_i exp(sigma_i W_i + (mu_i - sigma_i^2 / 2) tau).
    A transposed factor, or z taken in another order, moves R by O(sigma sqrt tau)."""
    tau, seed, ell, off, M = 0.25, 13, 3, 11, 500
    pf = _book(Q, (Q, 0, 0), [1] * Q, [100.0] * Q, [1.0] * Q, [1.0] * Q, list(range(Q)), tau)
    L = _chol(Q)
    O = Oracle("f64").new(1, 32, 2, 0.0, True)
    O.portfolio_init(pf, seed, chol=L)
    R, const = O.outer_states(ell, M, off)
    keys = (np.uint64(ell) << np.uint64(46)) | (np.uint64(off) + np.arange(M, dtype=np.uint64))
    z = rng_np.stream_normals(seed, keys, Q)
    W = math.sqrt(tau) * z @ L.T
    S0, mu, sig = pf["S0"], pf["mu"], pf["sigma"]
    R_np = S0 * np.exp(sig * W + (mu - 0.5 * sig * sig) * tau)
    np.testing.assert_allclose(R, R_np, rtol=1e-13, atol=0)
    # The transposed factor would not match (the check has teeth).
    W_t = math.sqrt(tau) * z @ L
    assert not np.allclose(S0 * np.exp(sig * W_t + (mu - 0.5 * sig * sig) * tau), R, rtol=1e-6)
    assert np.all(const == -pf["max_loss"])  # no control variate: constant = -max_loss


def test_cholesky_outer_draw_covariance():
    """Log-returns over 10^5 outer samples have covariance tau diag(sigma) L L^T
    diag(sigma), within 4 standard errors per entry (Var of a sample covariance
    of jointly normal x, y: (S_xx S_yy + S_xy^2) / n)."""
    Q, tau, n = 3, 0.25, 100_000
    pf = _book(Q, (Q, 0, 0), [1] * Q, [100.0] * Q, [1.0] * Q, [1.0] * Q, list(range(Q)), tau)
    L = _chol(Q)
    O = Oracle("f64").new(1, 32, 2, 0.0, True)
    O.portfolio_init(pf, 21, chol=L)
    R, _ = O.outer_states(0, n, 0)
    x = np.log(R / pf["S0"])
    S = tau * np.outer(pf["sigma"], pf["sigma"]) * (L @ L.T)
    C = np.cov(x.T)
    se = np.sqrt((np.outer(np.diag(S), np.diag(S)) + S * S) / n)
    assert np.all(np.abs(C - S) < 4 * se), (C, S, se)
    mean = x.mean(axis=0)
    drift = (pf["mu"] - 0.5 * pf["sigma"] ** 2) * tau
    assert np.all(np.abs(mean - drift) < 4 * np.sqrt(np.diag(S) / n))


def _bs_value(typ, S, K, r, t, sig):
    """Black-Scholes value: calls (1), puts (-1) and cash-or-nothing digitals
    paying 1 (2: S_T > K, -2: S_T < K)."""
    d1 = (np.log(S / K) + (r + 0.5 * sig * sig) * t) / (sig * np.sqrt(t))
    d2 = d1 - sig * np.sqrt(t)
    if abs(typ) == 2:
        return np.exp(-r * t) * _ncdf(np.sign(typ) * d2)
    return typ * S * _ncdf(typ * d1) - typ * K * np.exp(-r * t) * _ncdf(typ * d2)


@pytest.mark.parametrize("types", [(2, -2), (1, -1)], ids=["digital", "vanilla"])
def test_exact_inner_mean_closed_form(types):
    """Exact-GBM inner sampling of two options (a digital call and put, or a
    call and a put) on correlated stocks, all options in every inner sample, no
    control variate: given the outer draw R, the inner mean has expectation
    sum_i w_i (V_i(S0, T) - e^{-r tau} V_i(R, T - tau)) with V the option's
    Black-Scholes value.  Checked in aggregate over M outer samples (4
    standard errors) and, for the bounded digital terms (each within
    +-sum |w_i| e^{-r T_i}), per outer sample by Hoeffding's bound: a digital on
    the wrong side of the strike, the real-world drift in the inner paths or a
    missing discount fails."""
    Q, tau, r, N, M = 2, 0.25, 0.05, 4096, 3000
    K = [101.0, 97.0]
    T = [0.75, 1.25]
    w = [1.3, -0.7]
    pf = _book(Q, (0, 2, 0), list(types), K, T, w, [0, 1], tau, r=r)
    L = _chol(Q)
    O = Oracle("f64").new(1, N, 2, 0.0, True, moments=1)
    O.portfolio_init(pf, 5, chol=L)
    o = O.sample(0, M, keying=1, outer_offset=0, trace=True)
    assert np.all(o["nf"] == N)
    R, _ = O.outer_states(0, M, 0)
    S0, mu, sig = pf["S0"], pf["mu"], pf["sigma"]

    def closed_form(tps, drift):
        c = np.zeros(M)
        for typ, k, t, wi, s in zip(tps, K, T, w, [0, 1]):
            v0 = _bs_value(typ, S0[s], k, r, t, sig[s])
            # the value at tau of an option on a stock drifting at `drift`
            # (the stock's forward moved by e^{(drift - r)(T - tau)})
            vt = np.exp(-r * tau) * _bs_value(typ, R[:, s] * np.exp((drift[s] - r) * (t - tau)), k, r,
                                              t - tau, sig[s])
            c += wi * (v0 - vt)
        return c

    rr = np.full(Q, r)
    dev = o["phi_f"] - closed_form(types, rr)
    se = dev.std(ddof=1) / math.sqrt(M)
    assert abs(dev.mean()) < 4 * se, (dev.mean(), se)
    if abs(types[0]) == 2:
        B = sum(abs(wi) * math.exp(-r * t) for wi, t in zip(w, T))
        assert np.max(np.abs(dev)) < B * math.sqrt(2 * math.log(2 * M / 1e-9) / N)
        # The check has teeth: the digitals' sides swapped move the mean by far more.
        assert abs((o["phi_f"] - closed_form(types[::-1], rr)).mean()) > 20 * se
    else:
        # ... and the real-world drift in the inner paths (mu instead of r).
        assert abs((o["phi_f"] - closed_form(types, mu)).mean()) > 10 * se
    # ... and so would a loss scaled by 0.9 (a dropped term of that size).
    assert abs(o["phi_f"].mean() - 0.9 * closed_form(types, rr).mean()) > 10 * se


def _capped_counts(seed, ell_key, m, start, base, cap_log2, tag0, tmp=9.0):
    """get_samples_count under the capped 3-sigma rule (tmp = 9): from base,
    test N inner samples of the shift model X = Y + Z (draw group tag0 + round)
    and double while 9 / delta^2 > N; the step to the cap (max <= 2 N) is taken
    without a test.  Returns (count, smallest relative decision margin)."""
    key1 = (np.uint64(ell_key) << np.uint64(46)) | np.uint64(m)
    a, b = rng_np.threefry2x64_20(0, 0, seed, key1)
    Y, _ = rng_np.boxmuller(a, b)
    mx = base << cap_log2
    cur, rnd, margin = base, 0, np.inf
    while True:
        if mx <= 2 * cur:
            return mx, margin
        u = np.arange(cur // 2, dtype=np.uint64)
        ctr1 = (np.uint64(tag0 + rnd) << np.uint64(40)) | u
        a, b = rng_np.threefry2x64_20(np.uint64(0), ctr1, seed, key1)
        x, y = rng_np.boxmuller(a, b)
        d = Y + np.stack([x, y], axis=1).reshape(-1)
        s1, s2 = d.sum(), (d * d).sum()
        var = s2 / (cur - 1) - s1 * s1 / (cur * (cur - 1))
        dist = start + s1 / cur
        new = tmp / (dist * dist / var)
        margin = min(margin, abs(new - cur) / cur)
        rnd += 1
        if new <= cur:
            return cur, margin
        cur *= 2


def test_capped_three_sigma_counts_against_numpy_restatement():
    """N_f and N_c of the oracle's capped 3-sigma rule under the batched keying
    (fine tests tags 1 + a at level l, coarse tests tags 33 + a at level l - 1's
    counts) against the restatement above, on every outer sample that refined
    and on 40 that did not; the set includes samples taken to the cap by the
    forced last step."""
    seed, ell, M, cap = 3, 2, 4000, 6
    O = Oracle("f64").new(1, 32, 2, 0.0, True, adapt_rule=1, cap_log2=cap)
    O.model_init(1, 0, 0.5, 2.0, seed)  # shift model, indicator, threshold 2
    o = O.sample(ell, M, keying=1, outer_offset=0, trace=True)
    base_f, base_c = 32 << ell, 32 << (ell - 1)
    refined = np.nonzero((o["nf"] > base_f) | (o["nc"] > base_c))[0]
    others = np.setdiff1d(np.arange(M), refined)[:40]
    assert len(refined) >= 20
    assert np.any(o["nf"] == base_f << cap) and np.any((o["nf"] > base_f) & (o["nf"] < base_f << cap))
    checked = 0
    for m in np.concatenate([refined, others]):
        nf, mf = _capped_counts(seed, ell, int(m), -2.0, base_f, cap, 1)
        nc, mc = _capped_counts(seed, ell, int(m), -2.0, base_c, cap, 33)
        if min(mf, mc) < 1e-9:  # a decision on a rounding knife edge: no verdict
            continue
        assert (nf, nc) == (int(o["nf"][m]), int(o["nc"][m])), m
        checked += 1
    assert checked >= len(refined) + 35
    # The check has teeth: a 2-sigma rule (tmp = 4) gives other counts on
    # many of these samples.
    diff = sum(_capped_counts(seed, ell, int(m), -2.0, base_f, cap, 1, tmp=4.0)[0] != int(o["nf"][m])
               for m in refined[:20])
    assert diff >= 5
