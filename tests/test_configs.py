"""Generators of the benchmark books (CPU): the batched-risk books with an Asian
share, and the geometric-average proxy that serves their pilot thresholds."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cfg(nm):
    return importlib.import_module(nm.__name__ + ".configs")


def test_batch_books_asian_share(cfg):
    """20% of each book's options are Asian calls / puts, all of them (and
    nothing else) in the Unbiased method's span; each span is sorted by
    descending probability |w|; the tables and the Cholesky factor fit."""
    books = cfg.batch_books(3, seed=7, n_pilot=40, workers=1, asian_frac=0.2)
    for a, L, sc, q in books:
        P = len(a["w"])
        n1, n2 = int(a["counts"][1]), int(a["counts"][2])
        assert a["counts"][0] == 0 and n1 + n2 == P
        asian = np.abs(a["type"]) == 3
        assert asian.sum() == round(0.2 * P) == n2
        assert asian[n1:].all() and not asian[:n1].any()
        for lo, hi in ((0, n1), (n1, P)):
            assert np.all(np.diff(a["prob"][lo:hi]) <= 0)
        np.testing.assert_array_equal(a["prob"], np.abs(a["w"]))
        d = len(a["S0"])
        assert L.shape == (d, d) and np.all(a["stock_ind"] < d)
        assert q in (0.95, 0.99, 0.995) and np.isfinite(sc["max_loss"])


def test_batch_books_without_asians(cfg):
    """asian_frac = 0: every option is priced by the Exact method."""
    (a, L, sc, q), = cfg.batch_books(1, seed=7, n_pilot=40, workers=1, asian_frac=0.0)
    assert a["counts"][2] == 0 and not np.any(np.abs(a["type"]) == 3)


def test_geometric_asian_proxy(cfg):
    """The Asian proxy is the geometric-average Black-Scholes value: put-call
    parity with carry b = (r - sigma^2 / 6) / 2, i.e. C - P = e^(-r t)
    (S e^(b t) - K); it lies below the European call for r >= 0; as sigma -> 0
    the call tends to the discounted intrinsic value of S e^(b t)."""
    S, K, r, t = 100.0, np.array([80.0, 100.0, 120.0]), 0.05, 1.5
    for sig in (0.1, 0.3):
        c = cfg.bs_value(np.full(3, 3), S, K, r, t, sig)
        p = cfg.bs_value(np.full(3, -3), S, K, r, t, sig)
        b = 0.5 * (r - sig * sig / 6.0)
        np.testing.assert_allclose(c - p, np.exp(-r * t) * (S * np.exp(b * t) - K), rtol=1e-12, atol=1e-12)
        e = cfg.bs_value(np.full(3, 1), S, K, r, t, sig)
        assert np.all(c < e)
    c0 = cfg.bs_value(np.full(3, 3), S, K, r, t, 1e-6)
    np.testing.assert_allclose(c0, np.exp(-r * t) * np.maximum(S * np.exp(0.5 * r * t) - K, 0), atol=1e-4)
