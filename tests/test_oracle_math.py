"""The oracle pinned by what the mathematics fixes (no reference build needed):
published known-answer vectors, closed forms, brute force and invariants.  Each
check is chosen so that a plausible slip (a dropped term, a swapped operand, a
wrong rotation or index) fails it."""
import math

import numpy as np
import pytest
from scipy import special, stats

from oracle import Oracle


@pytest.fixture(scope="module")
def O():
    return Oracle("f64")


# Random123's published known-answer vectors for threefry2x64 with 20 rounds.
KAT = [
    ((0, 0), (0, 0), (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)),
    ((2**64 - 1, 2**64 - 1), (2**64 - 1, 2**64 - 1), (0xE02CB7C4D95D277A, 0xD06633D0893B8B68)),
    ((0x243F6A8885A308D3, 0x13198A2E03707344), (0xA4093822299F31D0, 0x082EFA98EC4E6C89),
     (0x263C7D30BB0F0AF1, 0x56BE8361D3311526)),
]


def test_threefry_published_kat(O):
    for ctr, key, out in KAT:
        assert tuple(O.threefry(list(ctr), list(key))) == out


def test_threefry_distinct_blocks(O):
    seen = {tuple(O.threefry([i, 0], [1, 2])) for i in range(4096)}
    assert len(seen) == 4096


def test_uniform_closed_forms(O):
    assert O.u01_f64(0) == 2.0 ** -65
    assert O.u01_f64(2**64 - 1) == 1.0
    assert O.u01_f64(2**63) == 0.5 + 2.0 ** -65
    assert O.uneg11_f64(0) == 2.0 ** -64
    assert O.uneg11_f64(2**63 - 1) == 1.0
    assert O.uneg11_f64(2**63) == -1.0
    assert O.uneg11_f64(2**64 - 1) == -(2.0 ** -63) + 2.0 ** -64
    assert np.float32(O.u01_f32(0)) == np.float32(2.0 ** -65)


def test_boxmuller_formula(O):
    rs = np.random.RandomState(1)
    for _ in range(200):
        u0, u1 = (int(x) for x in rs.randint(0, 2**63, 2, dtype=np.int64))
        u0 |= int(rs.randint(2)) << 63
        x, y = O.boxmuller(u0, u1)
        a = O.uneg11_f64(u0)
        r = math.sqrt(-2.0 * math.log(O.u01_f64(u1)))
        assert x == pytest.approx(math.sin(3.1415926535897932 * a) * r, rel=1e-15, abs=1e-300)
        assert y == pytest.approx(math.cos(3.1415926535897932 * a) * r, rel=1e-15, abs=1e-300)


def test_stream_consumption_order(O):
    """rand() hands out cur[1] then cur[0] of block ctr = 0, 1, ...; randn()
    builds boxmuller(u0 = second word taken, u1 = first word taken) and caches y."""
    seed, core = 9, 4
    u, _ = O.stream(seed, core, [0] * 6)
    words = []
    for b in range(3):
        blk = O.threefry([b, 0], [seed, core])
        words += [blk[1], blk[0]]
    assert [int(x) for x in u] == words
    _, r = O.stream(seed, core, [2, 2, 2])
    x0, y0 = O.boxmuller(words[1], words[0])
    x1, _ = O.boxmuller(words[3], words[2])
    assert list(r) == [x0, y0, x1]
    # A uniform between normals takes the next word and leaves the cached y in
    # place; the next pair then straddles two blocks (words 3 and 4).
    _, r = O.stream(seed, core, [2, 1, 2, 2])
    xa, ya = O.boxmuller(words[1], words[0])
    xb, _ = O.boxmuller(words[4], words[3])
    assert list(r) == [xa, O.u01_f64(words[2]), ya, xb]


def test_rand_int_cdf_brute_force(O):
    rs = np.random.RandomState(3)
    for n in (1, 2, 3, 5, 8, 9, 31, 100, 1000):
        p = np.sort(rs.rand(n) + 1e-3)[::-1]
        cdf = np.concatenate([[0.0], np.cumsum(p)[:-1]])
        for u in np.concatenate([rs.rand(64) * p.sum(), cdf[1:], [0.0, p.sum() * 0.999999]]):
            brute = int(np.searchsorted(cdf, u, side="right") - 1)
            assert O.rand_int_cdf(float(u), cdf) == brute


def test_rand_int_pdf_brute_force(O):
    rs = np.random.RandomState(4)
    for _ in range(200):
        pdf = rs.rand(3) * (rs.rand(3) > 0.3)
        if pdf.sum() == 0:
            pdf[1] = 0.5
        u = float(rs.rand() * pdf.sum() * 1.1)
        acc, ret = 0.0, 0
        for i, c in enumerate(pdf):
            if c <= 0:
                continue
            acc += c
            ret = i
            if u <= acc:
                break
        assert O.rand_int_pdf(u, pdf) == ret


def test_normal_cdf_and_inverse(O):
    for x in np.linspace(-6, 6, 41):
        assert O.norm_cdf(float(x)) == pytest.approx(stats.norm.cdf(x), rel=1e-14, abs=1e-300)
    for p in (1e-8, 1e-4, 0.01, 0.02425, 0.2, 0.5, 0.7, 0.99, 1 - 1e-6):
        # Acklam's rational approximation with single-precision coefficients.
        assert O.norm_cdf_inv(p) == pytest.approx(special.ndtri(p), rel=2e-6, abs=1e-7)


def test_black_scholes_identities(O):
    for S0, K, r, t, sig in ((100, 100, 0.05, 1.0, 0.2), (80, 120, 0.01, 0.3, 0.5), (130, 90, 0.0, 2.0, 0.15)):
        c, p = O.bs_price(1, S0, K, r, t, sig), O.bs_price(-1, S0, K, r, t, sig)
        assert c - p == pytest.approx(S0 - K * math.exp(-r * t), rel=1e-12)
        assert O.bs_delta(1, S0, K, r, t, sig) - O.bs_delta(-1, S0, K, r, t, sig) == pytest.approx(1.0)
        d1 = (math.log(S0 / K) + (r + sig * sig / 2) * t) / (sig * math.sqrt(t))
        assert c == pytest.approx(S0 * stats.norm.cdf(d1) - K * math.exp(-r * t) * stats.norm.cdf(d1 - sig * math.sqrt(t)), rel=1e-12)
        # Cash-or-nothing digitals (extension): call + put = e^{-rt}.
        dc, dp = O.bs_price(2, S0, K, r, t, sig), O.bs_price(-2, S0, K, r, t, sig)
        assert dc + dp == pytest.approx(math.exp(-r * t), rel=1e-12)
        h = 1e-4 * S0
        num = (O.bs_price(2, S0 + h, K, r, t, sig) - O.bs_price(2, S0 - h, K, r, t, sig)) / (2 * h)
        assert O.bs_delta(2, S0, K, r, t, sig) == pytest.approx(num, rel=1e-5)


def test_simulate_stock_dt_closed_forms(O):
    # sigma = 0: S (1 + mu dt)^N for fine paths, coarse paths with dt_c.
    N, mu, t = 16, 0.07, 0.5
    W, S = O.simulate_stock_dt(1, np.zeros(N), [100.0, 90.0, 100.0, 90.0], N, mu, 0.0, t, 2, 1, 4)
    assert W == 0.0
    assert S[0] == pytest.approx(100 * (1 + mu * t / N) ** N, rel=1e-14)
    assert S[2] == pytest.approx(100 * (1 + mu * t / (N / 4)) ** (N / 4), rel=1e-14)
    # Milstein step, brute force.
    rs = np.random.RandomState(0)
    z = rs.randn(8)
    sig, N = 0.3, 8
    dt = t / N
    W, S = O.simulate_stock_dt(0, z, [100.0, 0.0], N, mu, sig, t, 1, 1, 2)
    s = 100.0
    for n in range(N):
        dW = math.sqrt(dt) * z[n]
        s += s * (mu * dt + sig * dW + 0.5 * sig * sig * (dW * dW - dt))
    assert S[0] == pytest.approx(s, rel=1e-14)
    assert W == pytest.approx(math.sqrt(dt) * z.sum(), rel=1e-14)


def test_simulate_stock_dt_path_averages(O):
    """Path averages of the Asian extension: sigma = 0 gives the geometric
    sums (1/N) sum_k S (1 + mu dt)^k for the fine and the coarse paths; with
    noise, a brute-force Milstein recurrence with the averages summed in step
    order (fine: after each step; coarse: after each coarse_step-th), and the
    terminal values unchanged by the averaging."""
    N, mu, t = 16, 0.07, 0.5
    W, S, A = O.simulate_stock_dt_avg(1, np.zeros(N), [100.0, 90.0, 100.0, 90.0], N, mu, 0.0, t, 2, 1, 4)
    g = 1 + mu * t / N
    assert A[0] == pytest.approx(100 * sum(g ** k for k in range(1, N + 1)) / N, rel=1e-14)
    assert A[1] == pytest.approx(90 * sum(g ** k for k in range(1, N + 1)) / N, rel=1e-14)
    gc = 1 + mu * t / (N / 4)
    assert A[2] == pytest.approx(100 * sum(gc ** k for k in range(1, N // 4 + 1)) / (N // 4), rel=1e-14)
    rs = np.random.RandomState(3)
    sig, N, cs = 0.35, 12, 3
    z = rs.randn(N)
    W, S, A = O.simulate_stock_dt_avg(1, z, [100.0, 80.0], N, mu, sig, t, 1, 1, cs)
    W2, S2 = O.simulate_stock_dt(1, z, [100.0, 80.0], N, mu, sig, t, 1, 1, cs)
    assert list(S) == list(S2) and W == W2
    dt, dtc = t / N, t / (N // cs)
    s, sc, a, ac, dwc = 100.0, 80.0, 0.0, 0.0, 0.0
    for n in range(N):
        dW = math.sqrt(dt) * z[n]
        s += s * (mu * dt + sig * dW + 0.5 * sig * sig * (dW * dW - dt))
        a += s
        dwc += dW
        if (n + 1) % cs == 0:
            sc += sc * (mu * dtc + sig * dwc + 0.5 * sig * sig * (dwc * dwc - dtc))
            ac += sc
            dwc = 0.0
    assert A[0] == pytest.approx(a / N, rel=1e-14)
    assert A[1] == pytest.approx(ac / (N // cs), rel=1e-14)


def test_asian_inner_mean_closed_form(O):
    """An Asian call (type 3) priced by the Unbiased method with sigma = 0:
    both paths grow deterministically, and the randomised-level estimator's
    mean is the continuously averaged payoff difference
    w e^{-rT} (A(S0 e^{r tau}) - A(S0 e^{mu tau})) with A(x) = x (e^{r s} - 1)/(r s) - K,
    s = T - tau.  The terminal-value payoff, or a dropped coarse term, misses
    it by far more than the tolerance (r s = 0.6 separates the average from
    the terminal value by a third)."""
    r, mu, tau, T, K, w, S0 = 0.3, 0.1, 0.1, 2.1, 50.0, 1.0, 100.0
    pf = dict(S0=[S0], mu=[mu], sigma=[0.0], counts=np.array([0, 0, 1], np.uint32), strike=[K],
              maturity=[T], start=[0.0], w=[w], delta_est=[0.0], stock_ind=np.array([0], np.uint32),
              type=np.array([3], np.int32), prob=[1.0], W_est=[1.0], r=r, rho=0.0, max_loss=0.0,
              tau=tau, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5, phi_type=2, phi_size=1, subsample=1,
              use_delta_cv=False, use_coarse_delta_cv=False, use_antithetic_risk=True, use_milstein=True)
    O.new(1, 4096, 2, 0.0, True, moments=1)
    O.portfolio_init(pf, 5)
    o = O.sample(0, 16)
    s_ = T - tau
    avg = lambda x: x * (math.exp(r * s_) - 1) / (r * s_)
    exact = w * math.exp(-r * T) * ((avg(S0 * math.exp(r * tau)) - K) - (avg(S0 * math.exp(mu * tau)) - K))
    assert o["fsums"][0] / 16 == pytest.approx(exact, rel=0.02)
    terminal = w * math.exp(-r * T) * (S0 * math.exp(r * tau) - S0 * math.exp(mu * tau)) * math.exp(r * s_)
    assert abs(terminal - exact) > 0.2 * abs(exact)


def test_sample_counts(O):
    """get_samples_count's N_base / N_max: reference rule N0 (beta/2)^l, N0 beta^l;
    capped rule N0 beta^l and << cap."""
    O.new(1, 32, 4, 1.5, True)
    O.model_init(0, 0, 0.02, 0.0, 1)
    assert [O.samples_count(l) for l in range(4)] == [(32, 32), (64, 128), (128, 512), (256, 2048)]
    O.new(1, 32, 4, 0.0, True)
    O.model_init(0, 0, 0.02, 0.0, 1)
    assert [O.samples_count(l) for l in range(3)] == [(32, 32), (128, 128), (512, 512)]
    O.new(1, 32, 2, 0.0, True, adapt_rule=1, cap_log2=6)
    O.model_init(1, 0, 0.5, 2.0, 1)
    assert [O.samples_count(l) for l in range(3)] == [(32, 2048), (64, 4096), (128, 8192)]


def test_antithetic_level_has_zero_mean_difference_when_nonadaptive(O):
    """With the antithetic estimator E[phi_f - phi_c] -> bias differences; for
    the shift model at large N both are close: the level-l differences are
    mostly exact zeros (same draws, same sign)."""
    O.new(1, 32, 2, 0.0, True, adapt_rule=1, cap_log2=0)
    O.model_init(1, 0, 0.5, 2.0, 1)
    o = O.sample(4, 2000, keying=1, outer_offset=0, trace=True)
    d = o["phi_f"] - o["phi_c"]
    assert (d == 0).mean() > 0.95
    assert np.array_equal(o["nf"], np.full(2000, 512, np.uint64))
    assert np.array_equal(o["nc"], np.full(2000, 256, np.uint64))


def test_shift_model_closed_form():
    """P(E[X|Y] > 2) = 1 - Phi(2) with X = Y + Z (batched keying, FP64)."""
    O = Oracle("f64").new(1, 32, 2, 0.0, True, adapt_rule=1, cap_log2=6)
    O.model_init(1, 0, 0.5, 2.0, 17)
    est, var = 0.0, 0.0
    for ell, M in ((0, 200000), (1, 60000), (2, 20000), (3, 6000)):
        o = O.sample(ell, M, keying=1, outer_offset=0)
        s = o["fsums"] if ell == 0 else o["diffsums"]
        m = s[0] / M
        est += m
        var += (s[1] / M - m * m) / M
        last = abs(m)
    exact = 1 - stats.norm.cdf(2.0)
    assert abs(est - exact) < 4 * math.sqrt(var) + 2 * last


def test_reference_model_closed_form():
    """Reference model: P(tau (Y^2 - 1) > max_loss) = 2 Phi(-|q|), q = Acklam(eta/2)."""
    O = Oracle("f64").new(4, 32, 4, 1.5, True)
    tau, eta = 0.02, 0.025
    q = O.norm_cdf_inv(eta / 2)
    O.model_init(0, 0, tau, tau * (q * q - 1), 2)
    est, var = 0.0, 0.0
    for ell, M in ((0, 100000), (1, 30000), (2, 10000), (3, 3000)):
        o = O.sample(ell, M)
        s = o["fsums"] if ell == 0 else o["diffsums"]
        m = s[0] / M
        est += m
        var += (s[1] / M - m * m) / M
        last = abs(m)
    exact = 2 * stats.norm.cdf(-abs(q))
    assert abs(est - exact) < 4 * math.sqrt(var) + 2 * last


def test_batched_keying_layout(O):
    """Batched keying: the outer normal of sample g at level l is the x of
    boxmuller from block {0,0} under key {seed, l<<46 | g}; with phi = Max the
    fine-only level exposes start + mean of the final draws (tag 65)."""
    O.new(1, 4, 2, 0.0, True)
    O.model_init(1, 1, 0.5, -50.0, 5)  # shift model, Max, threshold -50: phi = 50 + mean
    o = O.sample(0, 3, keying=1, outer_offset=7, trace=True)
    for m in range(3):
        key = [5, (0 << 46) | (7 + m)]
        blk = O.threefry([0, 0], key)
        Y, _ = O.boxmuller(blk[0], blk[1])
        zs = []
        for u in range(2):
            b = O.threefry([0, (65 << 40) | u], key)
            zs += list(O.boxmuller(b[0], b[1]))
        mean = sum(Y + z for z in zs) / 4
        assert o["phi_f"][m] == pytest.approx(50 + mean, rel=1e-15)
