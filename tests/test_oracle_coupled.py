"""Pins of the oracle's level-coupled paths (an extension: the batched-risk
workload's arithmetic-average Asians on Milstein paths of dt_h0inv 2^floor(l/2)
steps at level l, fine and coarse values of an antithetic final from the same
increments).  Checked against a closed form (sigma = 0) and against a numpy
restatement built from the Threefry words (rng_np) for single inner samples,
so that a wrong step schedule, a dropped coarse coupling, averages over the
wrong steps or a misplaced draw fails."""
import math

import numpy as np
import pytest

from oracle import Oracle

import rng_np

TAG_FINAL = 65


def _book(types, strikes, mats, w, sigma=0.3, mu=0.07, cv=False):
    n = len(types)
    p = np.abs(np.array(w, float))
    order = np.argsort(-p, kind="stable")
    return dict(S0=np.array([100.0]), mu=np.array([mu]), sigma=np.array([sigma]),
                counts=np.array([0, 0, n], np.uint32), strike=np.array(strikes, float)[order],
                maturity=np.array(mats, float)[order], start=np.zeros(n), w=np.array(w, float)[order],
                delta_est=np.zeros(n), stock_ind=np.zeros(n, np.uint32),
                type=np.array(types, np.int32)[order], prob=p[order], W_est=np.ones(n), r=0.05, rho=0.0,
                max_loss=0.0, tau=0.1, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5, phi_type=2, phi_size=1,
                subsample=1, use_delta_cv=cv, use_coarse_delta_cv=False, use_antithetic_risk=True,
                use_milstein=True)


def test_coupled_paths_sigma_zero_closed_form():
    """sigma = 0: every path grows by (1 + r dt) per step from its value at tau
    (S0 e^{r tau} for the S0 paths, R = S0 e^{mu tau} for the outer sample's),
    so an Asian call on n steps pays max(X (1/n) sum_k (1 + r dt)^k - K, 0).  At
    level 2 the fine values take n = 16 steps and the coarse values n = 8 (the
    level-1 schedule), from one inner sample per group (N0 = 1, beta = 2, no
    refinement): phi_f and phi_c (Moments, the inner means) equal the closed
    forms with 16 and 8 steps."""
    r, tau, T, K, w = 0.05, 0.1, 1.1, 95.0, 1.0
    pf = _book([3], [K], [T], [w], sigma=0.0, mu=0.07)
    O = Oracle("f64").new(1, 1, 2, 0.0, True, moments=1)
    O.portfolio_init(pf, 3)
    O.set_coupled(True)

    def value(n):
        dt = (T - tau) / n
        g = sum((1 + r * dt) ** k for k in range(1, n + 1)) / n
        x0, xr = 100.0 * math.exp(r * tau), 100.0 * math.exp(0.07 * tau)
        return w * math.exp(-r * T) * (max(x0 * g - K, 0.0) - max(xr * g - K, 0.0))

    o = O.sample(2, 5, keying=1, outer_offset=0, trace=True)
    assert np.all(o["nf"] == 4) and np.all(o["nc"] == 2)
    np.testing.assert_allclose(o["phi_f"], value(16), rtol=1e-12)
    np.testing.assert_allclose(o["phi_c"], value(8), rtol=1e-12)
    # The schedule has teeth: 8 and 16 steps give different values here.
    assert abs(value(16) - value(8)) > 1e-3 * abs(value(8))
    # Level 1: fine and coarse both on 8 steps (floor(1/2) = floor(0/2)).
    o1 = O.sample(1, 3, keying=1, outer_offset=0, trace=True)
    np.testing.assert_allclose(o1["phi_f"], value(8), rtol=1e-12)
    np.testing.assert_allclose(o1["phi_c"], value(8), rtol=1e-12)


def _coupled_sample(pf, seed, ell, m, i, n_f, cs, two):
    """One inner sample of draw group TAG_FINAL of outer sample m at level ell,
    restated from the stream's words: word 0 the method uniform, word 1 the
    option uniform, then normals (Box-Muller of blocks 1, 2, ...): the first
    for the S0 paths to tau, the next n_f for the Milstein steps.  Returns
    (fine, coarse) values divided by the option's probability."""
    key1 = np.uint64((ell << 46) | m)
    c1 = np.uint64((TAG_FINAL << 40) | i)
    # outer draw: R = S0 exp(sigma W + (mu - sigma^2 / 2) tau), W = sqrt(tau) z0
    z_out = rng_np.stream_normals(seed, key1, 1)[0]
    S0, mu, sig = pf["S0"][0], pf["mu"][0], pf["sigma"][0]
    r, tau = pf["r"], pf["tau"]
    R = S0 * math.exp(sig * (math.sqrt(tau) * z_out[0]) + (mu - 0.5 * sig * sig) * tau)
    b0 = rng_np.threefry2x64_20(np.uint64(0), c1, seed, key1)
    u_opt = float(rng_np.u01(b0[0]))
    prob = pf["prob"]
    cdf = np.concatenate([[0.0], np.cumsum(prob)[:-1]])
    idx = int(np.max(np.nonzero(cdf <= u_opt * prob.sum())[0]))
    nblk = (n_f + 2) // 2 + 1
    z = []
    for j in range(1, nblk + 1):
        a, b = rng_np.threefry2x64_20(np.uint64(j), c1, seed, key1)
        x, y = rng_np.boxmuller(a, b)
        z += [float(x), float(y)]
    typ, K, T, w = int(pf["type"][idx]), pf["strike"][idx], pf["maturity"][idx], pf["w"][idx]
    dW0 = math.sqrt(tau) * z[0]
    s1 = S0 * math.exp(sig * dW0 + (r - 0.5 * sig * sig) * tau)
    s2 = S0 * math.exp(sig * -dW0 + (r - 0.5 * sig * sig) * tau)
    S = [s1, R, s2, s1, R, s2]
    A = [0.0] * 6
    Tt = T - tau
    dt, sdt = Tt / n_f, math.sqrt(Tt / n_f)
    coarse = two and cs > 1
    dtc = Tt / (n_f // cs)
    dWc = 0.0
    for n in range(n_f):
        dW = sdt * z[1 + n]
        for k in range(3):
            S[k] += S[k] * (r * dt + sig * dW + 0.5 * (sig * sig) * (dW * dW - dt))
            A[k] += S[k]
        if coarse:
            dWc += dW
            if (n + 1) % cs == 0:
                for k in range(3):
                    S[k + 3] += S[k + 3] * (r * dtc + sig * dWc + 0.5 * (sig * sig) * (dWc * dWc - dtc))
                    A[k + 3] += S[k + 3]
                dWc = 0.0
    A = [a / n_f for a in A[:3]] + [a / (n_f // cs) for a in A[3:]]
    V = A if abs(typ) == 3 else S

    def payoff(x):
        mod = typ // 3 if abs(typ) == 3 else typ
        return max(mod * (x - K), 0.0)

    def derv(x):
        if abs(typ) == 3:
            return 0.0
        return w * (x / S0) * math.exp(-r * T) * (float(typ == 1) - float(x < K))

    vals = []
    for base in (0, 3) if coarse else (0,):
        v, vd = 0.0, 0.0
        for ic in (0, 2):
            v += 0.5 * (w * math.exp(-r * T) * (payoff(V[base + ic]) - payoff(V[base + 1])))
            if pf["use_delta_cv"]:
                vd += 0.5 * derv(S[base + ic])
        vals.append(v + (R - S0) * vd)
    if not coarse:
        vals.append(vals[0])
    p = prob[idx] / prob.sum()
    return vals[0] / p, vals[1] / p


@pytest.mark.parametrize("ell,cv", [(2, False), (3, True), (4, True)])
def test_coupled_paths_against_numpy_restatement(ell, cv):
    """Asian calls and puts and a European call (with the delta control variate
    on its own paths) in method 2; N0 = 1, beta = 2, no refinement, so that at
    level ell an outer sample draws N_f = 2^ell final inner samples and N_c =
    2^(ell-1): phi_f (Moments: the inner mean) is the mean of their fine values,
    phi_c the mean over the groups of N_c samples of their coarse values'
    means.  Each value restated from the Threefry words."""
    pf = _book([3, -3, 1, 3], [95.0, 105.0, 100.0, 110.0], [1.2, 0.8, 1.5, 2.0], [1.0, -0.7, 0.5, 0.3],
               sigma=0.3, mu=0.07, cv=cv)
    seed, M = 11, 4
    O = Oracle("f64").new(1, 1, 2, 0.0, True, moments=1)
    O.portfolio_init(pf, seed)
    O.set_coupled(True)
    o = O.sample(ell, M, keying=1, outer_offset=0, trace=True)
    nf, nc = 1 << ell, 1 << (ell - 1)
    assert np.all(o["nf"] == nf) and np.all(o["nc"] == nc)
    n_f = 8 << (ell // 2)
    cs = n_f // (8 << ((ell - 1) // 2))
    for m in range(M):
        vf, vc = zip(*[_coupled_sample(pf, seed, ell, m, i, n_f, cs, True) for i in range(nf)])
        # Big side is the fine one (N_f > N_c): phi_f = mean of the fine values,
        # phi_c = the mean over the groups of N_c of the coarse values' means.
        np.testing.assert_allclose(o["phi_f"][m], np.mean(vf), rtol=1e-11, atol=1e-13)
        groups = [np.mean(vc[g * nc:(g + 1) * nc]) for g in range(nf // nc)]
        np.testing.assert_allclose(o["phi_c"][m], np.mean(groups), rtol=1e-11, atol=1e-13)
        if cs > 1:  # a coupling with teeth: the coarse values differ from the fine ones
            assert np.max(np.abs(np.array(vf) - np.array(vc))) > 1e-6
