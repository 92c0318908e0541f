#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the reference's own OpenMP
build (oracle/_ref/, compiled in place from the reference sources by
oracle/Makefile).  Run in a container where the reference tree exists:

    make -C oracle ref && python3 tests/golden/make_goldens.py

Every fixture is data: inputs and the reference's outputs for them.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
from oracle import Ref  # noqa: E402


def u64s(a):
    return [str(int(x)) for x in a]


def rng_fixture(R):
    out = {"threefry": [], "streams": [], "u01": [], "boxmuller": []}
    cases = [([0, 0], [0, 0]), ([0xFFFFFFFFFFFFFFFF] * 2, [0xFFFFFFFFFFFFFFFF] * 2),
             ([0x243F6A8885A308D3, 0x13198A2E03707344], [0xA4093822299F31D0, 0x082EFA98EC4E6C89]),
             ([5, 0], [1, 7]), ([2**63 + 11, 3], [0, 2**40 + 5])]
    rs = np.random.RandomState(2024)
    for _ in range(8):
        cases.append(([int(x) for x in rs.randint(0, 2**63, 2, dtype=np.int64)],
                      [int(x) for x in rs.randint(0, 2**63, 2, dtype=np.int64)]))
    for ctr, key in cases:
        out["threefry"].append({"ctr": u64s(ctr), "key": u64s(key), "out": u64s(R.threefry(ctr, key))})
    for seed, core in ((0, 0), (1, 0), (1, 7), (12345, 65535)):
        kinds = rs.randint(0, 3, 64).astype(np.int32)
        u, r = R.stream(seed, core, kinds)
        out["streams"].append({"seed": seed, "core": core, "kinds": kinds.tolist(), "u": u64s(u),
                               "r": r.tolist()})
    for v in [0, 1, 2**63 - 1, 2**63, 2**64 - 1, 2**52 + 3, 123456789123456789]:
        out["u01"].append({"u": str(v), "u01_f64": R.u01_f64(v), "uneg11_f64": R.uneg11_f64(v),
                           "u01_f32": float(R.u01_f32(v))})
    for _ in range(16):
        a, b = (int(x) for x in rs.randint(0, 2**63, 2, dtype=np.int64))
        a ^= int(rs.randint(0, 2)) << 63
        out["boxmuller"].append({"u0": str(a), "u1": str(b), "out": list(R.boxmuller(a, b))})
    return out


def util_fixture(R):
    rs = np.random.RandomState(7)
    out = {"cdf": [], "pdf": [], "norm_cdf_inv": [], "norm_cdf": [], "bs": [], "ssdt": []}
    for n in (1, 2, 3, 7, 64, 100):
        p = np.sort(rs.rand(n))[::-1]
        cdf = np.concatenate([[0.0], np.cumsum(p)[:-1]])
        us = (rs.rand(12) * p.sum()).tolist() + [0.0, float(cdf[-1]), float(p.sum())]
        out["cdf"].append({"cdf": cdf.tolist(), "u": us, "idx": [R.rand_int_cdf(u, cdf) for u in us]})
    for pdf in ([0.3, 0.5, 0.2], [0.0, 0.4, 0.6], [0.5, 0.0, 0.5], [0.0, 0.0, 1.0]):
        us = [0.0, 0.1, 0.3, 0.31, 0.5, 0.79, 0.8, 0.81, 1.0]
        out["pdf"].append({"pdf": pdf, "u": us, "idx": [R.rand_int_pdf(u, pdf) for u in us]})
    for p in (1e-9, 1e-3, 0.0125, 0.02425, 0.3, 0.5, 0.8, 0.97575, 0.99, 1 - 1e-7):
        out["norm_cdf_inv"].append([p, R.norm_cdf_inv(p)])
    for x in (-8.0, -2.0, -0.3, 0.0, 0.7, 2.0, 5.0):
        out["norm_cdf"].append([x, R.norm_cdf(x)])
    for typ in (1, -1):
        for S0, K, r, t, sig in ((100, 100, 0.05, 1.0, 0.2), (95, 110, 0.03, 0.25, 0.4),
                                 (120, 90, 0.0, 2.0, 0.1)):
            out["bs"].append({"type": typ, "args": [S0, K, r, t, sig],
                              "price": R.bs_price(typ, S0, K, r, t, sig),
                              "delta": R.bs_delta(typ, S0, K, r, t, sig)})
    for with_coarse, N, milstein, step in ((0, 8, 1, 2), (1, 16, 1, 4), (1, 8, 0, 2)):
        normals = rs.randn(N).tolist()
        S = [100.0, 101.0, 99.0, 100.0, 101.0, 99.0]
        W, Sout = R.simulate_stock_dt(with_coarse, normals, S, N, 0.05, 0.3, 0.75, 3, milstein, step)
        out["ssdt"].append({"with_coarse": with_coarse, "N": N, "milstein": milstein,
                            "coarse_step": step, "normals": normals, "S": S, "W": W,
                            "S_out": Sout.tolist()})
    return out


def model_fixture(R):
    """Consecutive level calls of model_sample (state persists across calls)."""
    out = []
    tau, eta = 0.02, 0.025
    ml = tau * (R.norm_cdf_inv(eta / 2) ** 2 - 1)
    runs = [dict(cores=1, N0=32, beta=4, r=1.5, anti=True, phi=0, seed=0),
            dict(cores=1, N0=32, beta=4, r=0.0, anti=True, phi=1, seed=3),
            dict(cores=1, N0=16, beta=4, r=1.5, anti=False, phi=1, seed=5),
            dict(cores=8, N0=32, beta=4, r=1.5, anti=True, phi=0, seed=1),
            dict(cores=3, N0=32, beta=2, r=0.0, anti=True, phi=1, seed=9)]
    calls = [(0, 200), (1, 120), (2, 50), (0, 7), (3, 20), (1, 1), (1, 1), (1, 1)]
    for run in runs:
        R.init(run["cores"], run["N0"], run["beta"], run["r"], run["anti"])
        R.model_init(run["phi"], tau, ml, run["seed"])
        res = []
        for ell, M in calls:
            o = R.model_sample(ell, M)
            res.append({"ell": ell, "M": M, "fsums": o["fsums"].tolist(),
                        "diffsums": o["diffsums"].tolist(), "work": o["work"],
                        "stats": u64s(o["stats"])})
        out.append({"params": run, "tau": tau, "max_loss": ml, "calls": res})
    # Per-sample phi: single-thread M=1 calls expose one outer sample each.
    R.init(1, 32, 4, 1.5, True)
    R.model_init(1, tau, -0.05, 11)
    per = []
    for k in range(40):
        ell = [0, 1, 2][k % 3]
        o = R.model_sample(ell, 1)
        per.append({"ell": ell, "phi_f": o["fsums"][0], "dphi": o["diffsums"][0],
                    "stats": u64s(o["stats"])})
    return out, {"params": dict(cores=1, N0=32, beta=4, r=1.5, anti=True, phi=1, seed=11),
                 "tau": tau, "max_loss": -0.05, "samples": per}


def portfolio_fixture(R):
    """portfolio_gen_data output and level calls of portfolio_sample."""
    out = {}
    gen = R.portfolio_gen(60, 4, 1, [0.3, 0.5, 0.2])
    gen_j = {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in gen.items()}
    out["gen"] = {"P": 60, "Q": 4, "seed": 1, "probs": [0.3, 0.5, 0.2], "data": gen_j}
    runs = []
    for sub, cv, phi_type, phi_size, cores in ((1, True, 0, 1, 1), (1, True, 2, 2, 1),
                                               (-1, False, 1, 1, 1), (0, True, 3, 3, 1),
                                               (1, True, 0, 1, 4)):
        pf = dict(gen)
        pf.update(r=0.05, rho=0.2, max_loss=0.02, tau=0.02, subsample=sub, use_delta_cv=cv,
                  use_antithetic_risk=cv, use_milstein=True, use_coarse_delta_cv=False,
                  phi_type=phi_type, phi_size=phi_size, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5)
        R.init(cores, 8, 4, 1.5, True)
        start, delta = R.portfolio_set(pf)
        R.portfolio_init(3)
        calls = []
        for ell, M in ((0, 9), (1, 6), (2, 3)):
            o = R.portfolio_sample(ell, M)
            calls.append({"ell": ell, "M": M, "fsums": o["fsums"].tolist(),
                          "diffsums": o["diffsums"].tolist(), "work": o["work"],
                          "stats": u64s(o["stats"])})
        scal = {k: pf[k] for k in ("r", "rho", "max_loss", "tau", "subsample", "use_delta_cv",
                                   "use_antithetic_risk", "use_milstein", "use_coarse_delta_cv",
                                   "phi_type", "phi_size", "dt_h0inv", "dt_beta", "dt_prob_expo")}
        runs.append({"scalars": scal, "cores": cores, "N0": 8, "beta": 4, "r": 1.5,
                     "start": start.tolist(), "delta_est": delta.tolist(), "calls": calls})
    out["runs"] = runs
    # portfolio_sample_delta_unbiased: consecutive calls (persistent streams).
    deltas = []
    for cores in (1, 3):
        pf = dict(gen)
        pf.update(r=0.05, rho=0.2, max_loss=0.02, tau=0.02, subsample=1, use_delta_cv=True,
                  use_antithetic_risk=True, use_milstein=True, use_coarse_delta_cv=False,
                  phi_type=0, phi_size=1, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5)
        R.init(cores, 8, 4, 1.5, True)
        start, delta = R.portfolio_set(pf)
        R.portfolio_init(5)
        calls = []
        for idx, M in ((0, 40), (3, 25), (0, 7)):
            sums, work = R.portfolio_sample_delta(idx, M, 2)
            calls.append({"option_index": idx, "M": M, "sums": sums.tolist(), "work": work})
        deltas.append({"cores": cores, "seed": 5, "calls": calls})
    out["delta_unbiased"] = deltas
    # A second generated portfolio: other sizes and split, and all-equal
    # probabilities (the order of equal keys is the sort's own).
    extra = []
    for P, Q, seed, probs, up in ((200, 7, 42, [0.5, 0.25, 0.25], False),
                                  (64, 5, 3, [0.25, 0.5, 0.25], True)):
        g = R.portfolio_gen(P, Q, seed, probs, uniform_prob=up)
        extra.append({"P": P, "Q": Q, "seed": seed, "probs": probs, "uniform_prob": up,
                      "data": {k: (v.tolist() if hasattr(v, "tolist") else v)
                               for k, v in g.items()}})
    out["gen_more"] = extra
    return out


def mlmc_fixture(R):
    """mlmc_t::run call sequences for deterministic level samplers."""
    out = []
    L = 12
    cases = [dict(name="geometric", E=[2.0 ** -(l + 1) for l in range(L)],
                  V=[2.0 ** -(2 * l + 2) for l in range(L)], W=[2.0 ** l for l in range(L)],
                  Ef=[0.5] * L, Vf=[0.25] * L, tol=1e-3, M0=100),
             dict(name="nested-like", E=[0.02 * 2.0 ** -(1.0 * l) for l in range(L)],
                  V=[0.02 * 2.0 ** -(1.5 * l) for l in range(L)],
                  W=[32.0 * 2.0 ** l for l in range(L)], Ef=[0.0227] * L, Vf=[0.0222] * L,
                  tol=2e-3, M0=1024),
             # A level whose draws show no fine/coarse difference (V_3 = 0, as
             # a first batch of a rare-disagreement indicator level can): the
             # driver never samples it beyond its first M0.
             dict(name="zero-variance-level",
                  E=[0.0 if l == 3 else 0.02 * 2.0 ** -(1.0 * l) for l in range(L)],
                  V=[0.0 if l == 3 else 0.02 * 2.0 ** -(1.5 * l) for l in range(L)],
                  W=[32.0 * 2.0 ** l for l in range(L)], Ef=[0.0227] * L, Vf=[0.0222] * L,
                  tol=2e-4, M0=1024)]
    for cs in cases:
        ell, M, summ = R.mlmc_fake_run(cs["tol"], cs["M0"], cs["E"], cs["V"], cs["W"], cs["Ef"],
                                       cs["Vf"])
        out.append({"case": cs, "ell": ell.tolist(), "M": u64s(M), "summary": summ.tolist(),
                    "table": R.mlmc_last_table()})
    return out


def main():
    if not os.path.exists(os.path.join(HERE, "..", "..", "oracle", "_ref", "libref_f64.so")):
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference exists")
    for real in ("f64", "f32"):
        R = Ref(real)
        model, per = model_fixture(R)
        fx = {"real": real, "rng": rng_fixture(R), "util": util_fixture(R), "model": model,
              "model_per_sample": per, "portfolio": portfolio_fixture(R)}
        fx["layout"] = R.layout()
        if real == "f64":
            fx["mlmc"] = mlmc_fixture(R)
        with open(os.path.join(HERE, "reference_%s.json" % real), "w") as f:
            json.dump(fx, f, separators=(",", ":"))
        print("wrote reference_%s.json" % real)


if __name__ == "__main__":
    main()
