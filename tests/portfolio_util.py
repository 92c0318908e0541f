"""Small synthetic portfolios for tests (the same arrays go to the HIP library and
to the oracle)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def reference_generated(real="f64"):
    """The reference's portfolio_gen_data(60, 4, seed 1, probs 0.3/0.5/0.2)
    output (BS, Exact and Unbiased options), from the golden fixture."""
    with open(os.path.join(HERE, "golden", "reference_%s.json" % real)) as f:
        g = json.load(f)["portfolio"]["gen"]["data"]
    return {k: np.array(v) for k, v in g.items()}


def synthetic(P=40, Q=3, counts=(10, 30, 0), digitals=True, seed=0):
    """Random options per method, probabilities |w| sorted descending."""
    rs = np.random.RandomState(seed)
    d = dict(S0=rs.uniform(90, 110, Q), mu=rs.uniform(0.02, 0.1, Q), sigma=rs.uniform(0.1, 0.4, Q))
    cols = {k: [] for k in ("strike", "maturity", "w", "stock_ind", "type", "prob", "W_est")}
    for m, n in enumerate(counts):
        if n == 0:
            continue
        w = rs.randn(n)
        types = rs.choice([1, -1, 2, -2] if digitals else [1, -1], n)
        prob = np.abs(w) + 1e-3
        order = np.argsort(-prob, kind="stable")
        cols["strike"] += list(100 * rs.uniform(0.8, 1.2, n)[order])
        cols["maturity"] += list(rs.uniform(0.25, 2.0, n)[order])
        cols["w"] += list(w[order])
        cols["stock_ind"] += list(rs.randint(0, Q, n)[order])
        cols["type"] += list(types[order])
        cols["prob"] += list(prob[order])
        cols["W_est"] += [1.0] * n
    for k, v in cols.items():
        d[k] = np.array(v)
    d["counts"] = np.array(counts, np.uint32)
    return d


def to_library(nm, d, **scalars):
    pf = nm.Portfolio(d["S0"], d["mu"], d["sigma"], d["counts"], d["strike"], d["maturity"], d["w"],
                      d["stock_ind"], d["type"], d["prob"], d.get("W_est"), **scalars)
    start, delta = pf.comp_params()
    return pf, start, delta


def to_oracle(d, start, delta, **scalars):
    o = dict(d)
    o["start"], o["delta_est"] = start, delta
    if "W_est" not in o:
        o["W_est"] = np.ones(len(o["w"]))
    o.update(scalars)
    return o
