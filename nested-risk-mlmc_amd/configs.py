"""Portfolio workloads: correlated multi-asset option books and their loss
thresholds (generators the reference does not have; its own generator is
portfolio_gen_data).

Each generator returns ``(arrays, chol, scalars)``:

- ``arrays``: S0, mu, sigma (per stock), counts (per method BS / Exact /
  Unbiased), and strike, maturity, w, stock_ind, type, prob, W_est (per option,
  method-major, each method sorted by descending prob = |w|);
- ``chol``: the lower Cholesky factor of the asset correlation;
- ``scalars``: the portfolio_data scalars, including ``max_loss``, the pilot
  quantile of the loss.

The loss of an outer scenario is sum_i w_i (V_i(0) - e^{-r tau} V_i(tau)),
the value lost over the horizon tau (portfolio_sampler.cpp:186-219), where
V_i(tau) is the option's value at the scenario's stock price. For Exact-GBM
inner sampling, the inner mean given the scenario is exactly that sum with
Black-Scholes values. The pilot threshold is therefore the q-quantile of this
closed form over ``n_pilot`` scenarios drawn as the sampler draws them:
W = sqrt(tau) L z, R = S0 exp(sigma W + (mu - sigma^2/2) tau).
"""
import numpy as np
from scipy.special import erfc

SQRT2 = np.sqrt(2.0)


def _ncdf(x):
    return 0.5 * erfc(-x / SQRT2)


def bs_value(typ, S, K, r, t, sigma):
    """Black-Scholes value (calc_option_blackscholes, portfolio_sampler.cpp:140-150)
    of calls (1), puts (-1) and cash-or-nothing digitals paying 1 (2, -2).  For
    the arithmetic-average Asians (3, -3, averaging from now to t) it is the
    geometric-average closed form (sigma / sqrt 3, carry (r - sigma^2 / 6) / 2),
    a proxy that only serves the pilot threshold."""
    typ = np.asarray(typ)
    asian = np.abs(typ) == 3
    sig = np.where(asian, sigma / np.sqrt(3.0), sigma)
    b = np.where(asian, 0.5 * (r - sigma * sigma / 6.0), r)  # cost of carry
    d1 = (np.log(S / K) + (b + 0.5 * sig * sig) * t) / (sig * np.sqrt(t))
    d2 = d1 - sig * np.sqrt(t)
    mod = np.where(np.abs(typ) == 2, typ // 2, np.where(asian, typ // 3, typ)).astype(np.float64)
    vanilla = mod * S * np.exp((b - r) * t) * _ncdf(mod * d1) - mod * K * np.exp(-r * t) * _ncdf(mod * d2)
    digital = np.exp(-r * t) * _ncdf(mod * d2)
    return np.where(np.abs(typ) == 2, digital, vanilla)


def _book(rs, P, d, kinds, weights):
    """P options on d stocks: kinds = list of (type, fraction); weights 'normal'
    or 't3'.  Asian options (types 3 / -3) are priced by the Unbiased method
    (method 2, randomised-level Milstein paths), all others by Exact-GBM inner
    sampling (method 1); each method's options sorted by descending prob = |w|."""
    types = np.concatenate([np.full(int(round(f * P)), t, np.int32) for t, f in kinds])
    assert len(types) == P
    rs.shuffle(types)
    strike = 100.0 * rs.uniform(0.8, 1.2, P)
    maturity = rs.uniform(0.25, 2.0, P)
    stock_ind = rs.randint(0, d, P).astype(np.uint32)
    w = rs.standard_normal(P) if weights == "normal" else rs.standard_t(3, P)
    prob = np.abs(w)
    method = np.where(np.abs(types) == 3, 2, 1)
    order = np.lexsort((-prob, method))  # method-major, then descending prob
    n2 = int(np.sum(method == 2))
    return dict(strike=strike[order], maturity=maturity[order], w=w[order],
                stock_ind=stock_ind[order], type=types[order], prob=prob[order],
                W_est=np.ones(P), counts=np.array([0, P - n2, n2], np.uint32))


def conditional_loss(arrays, chol, scalars, z):
    """Closed-form E[loss | scenario] for the scenarios z [n x d] (standard normals)."""
    tau, r = scalars["tau"], scalars["r"]
    S0, mu, sig = arrays["S0"], arrays["mu"], arrays["sigma"]
    W = np.sqrt(tau) * (z @ chol.T)
    R = S0 * np.exp(sig * W + (mu - 0.5 * sig * sig) * tau)
    si = arrays["stock_ind"]
    typ, K, T, w = arrays["type"], arrays["strike"], arrays["maturity"], arrays["w"]
    v0 = bs_value(typ, S0[si], K, r, T, sig[si])
    out = np.empty(len(z))
    for k in range(len(z)):
        vt = bs_value(typ, R[k, si], K, r, T - tau, sig[si])
        out[k] = np.dot(w, v0 - np.exp(-r * tau) * vt)
    return out


def pilot_threshold(arrays, chol, scalars, q, n_pilot, seed):
    z = np.random.RandomState(seed).standard_normal((n_pilot, len(arrays["S0"])))
    return float(np.quantile(conditional_loss(arrays, chol, scalars, z), q))


def _scalars(tau, rho):
    return dict(r=0.05, rho=rho, tau=tau, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5, phi_type=0,
                phi_size=1, subsample=1, use_delta_cv=True, use_coarse_delta_cv=False,
                use_antithetic_risk=True, use_milstein=True, max_loss=0.0)


def small_portfolio(seed=3, n_pilot=4000, P=10_000, d=10):
    """d = 10 correlated GBM assets (S0 = 100, sigma ~ U(0.1, 0.4), drift 0.05,
    pairwise correlation 0.3 via its Cholesky factor), tau = 1/52; P = 10^4
    European calls and puts (50/50), strikes 100 U(0.8, 1.2), maturities
    U(0.25, 2), assets uniform, weights N(0, 1); Exact-GBM inner loss;
    sub-sampling with probability proportional to |w|; threshold = pilot 99%
    quantile of the loss."""
    rs = np.random.RandomState(seed)
    arrays = dict(S0=np.full(d, 100.0), mu=np.full(d, 0.05), sigma=rs.uniform(0.1, 0.4, d))
    arrays.update(_book(rs, P, d, [(1, 0.5), (-1, 0.5)], "normal"))
    C = 0.3 * np.ones((d, d)) + 0.7 * np.eye(d)
    chol = np.linalg.cholesky(C)
    sc = _scalars(1.0 / 52, 0.3)
    sc["max_loss"] = pilot_threshold(arrays, chol, sc, 0.99, n_pilot, seed + 1)
    return arrays, chol, sc


def large_portfolio(seed=4, n_pilot=1000, P=100_000, d=100, factors=5):
    """d = 100 assets with a factor-model correlation (5 factors, loadings
    U(0, 0.5), rows capped at systematic variance 0.95, idiosyncratic part to
    unit variance), S0 = 100, sigma ~
    U(0.1, 0.4), drift 0.05, tau = 1/52; P = 10^5 options: 40% calls, 40% puts,
    20% cash-or-nothing digitals, weights Student-t (nu = 3); Exact-GBM inner
    loss; threshold = pilot 99.5% quantile of the loss."""
    rs = np.random.RandomState(seed)
    B = rs.uniform(0.0, 0.5, (d, factors))
    # A row's systematic variance sum_k B_ik^2 can exceed 1 (up to 1.25): such
    # rows are scaled to 0.95 so that the idiosyncratic variance stays positive.
    ss = np.sum(B * B, axis=1)
    B *= np.sqrt(np.minimum(1.0, 0.95 / ss))[:, None]
    C = B @ B.T
    C[np.diag_indices(d)] = 1.0
    chol = np.linalg.cholesky(C)
    arrays = dict(S0=np.full(d, 100.0), mu=np.full(d, 0.05), sigma=rs.uniform(0.1, 0.4, d))
    arrays.update(_book(rs, P, d, [(1, 0.4), (-1, 0.4), (2, 0.1), (-2, 0.1)], "t3"))
    sc = _scalars(1.0 / 52, 0.0)
    sc["max_loss"] = pilot_threshold(arrays, chol, sc, 0.995, n_pilot, seed + 1)
    return arrays, chol, sc


def _batch_book(seed, k, n_pilot, asian_frac=0.2):
    rs = np.random.RandomState(seed * 1000 + k)
    d = (10, 50, 100)[rs.randint(3)]
    P = 10 * int(round(10 ** rs.uniform(3, 5) / 10))  # the 40/40/10/10 split stays whole
    q = (0.95, 0.99, 0.995)[rs.randint(3)]
    factors = min(5, d)
    B = rs.uniform(0.0, 0.5, (d, factors))
    ss = np.sum(B * B, axis=1)
    B *= np.sqrt(np.minimum(1.0, 0.95 / ss))[:, None]
    C = B @ B.T
    C[np.diag_indices(d)] = 1.0
    chol = np.linalg.cholesky(C)
    arrays = dict(S0=np.full(d, 100.0), mu=np.full(d, 0.05), sigma=rs.uniform(0.1, 0.4, d))
    a = asian_frac / 2  # P is a multiple of 10 and a of 0.05: the counts stay whole
    kinds = [(1, 0.4 - a), (-1, 0.4 - a), (2, 0.1), (-2, 0.1)] + ([(3, a), (-3, a)] if a else [])
    arrays.update(_book(rs, P, d, kinds, "t3"))
    sc = _scalars(1.0 / 52, 0.0)
    sc["max_loss"] = pilot_threshold(arrays, chol, sc, q, n_pilot, seed * 1000 + k + 500)
    return arrays, chol, sc, q


def _one_blas_thread():
    # Each generator process runs its small matrix products on one thread:
    # BLAS pools of every process spinning on the same cores are far slower.
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)


def batch_books(n=64, seed=5, n_pilot=200, workers=16, asian_frac=0.2):
    """Independent problems for batched risk.  Instance k (its own random
    stream, seed * 1000 + k) draws d in {10, 50, 100} assets (S0 = 100,
    sigma ~ U(0.1, 0.4), drift 0.05, factor-model correlation as in the large
    book with min(5, d) factors), P log-uniform in [10^3, 10^5] options
    (asian_frac arithmetic-average Asian calls and puts by the Unbiased
    method's randomised-level Milstein paths; 20% cash-or-nothing digitals and
    the rest calls and puts half and half, by Exact-GBM inner sampling;
    Student-t(3) weights),
    tau = 1/52, and a threshold at the pilot 95%, 99% or 99.5% quantile of its
    loss.  The instances are generated in up to `workers` processes.  Returns a
    list of (arrays, chol, scalars, q)."""
    import os
    from concurrent.futures import ProcessPoolExecutor
    w = max(1, min(workers, os.cpu_count() or 1, n))
    if w == 1:
        return [_batch_book(seed, k, n_pilot, asian_frac) for k in range(n)]
    with ProcessPoolExecutor(w, initializer=_one_blas_thread) as ex:
        return list(ex.map(_batch_book, [seed] * n, range(n), [n_pilot] * n, [asian_frac] * n))
