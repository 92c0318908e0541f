"""Why the MLMC estimate of the nested-risk model can sit several standard
errors away from the closed form P(L > 2) = Phi(-2) at eps = 1e-5.

The driver is the reference's mlmc_t::run (mlmc.hpp:290-337).  Its
get_optimal_Ml (mlmc.hpp:265-289) gives level l ceil(lambda sqrt(V_l / W_l))
samples, with V_l the sample variance of the level's draws so far: a level
whose first M0 differences are all zero gets V_l = 0 and is never sampled
again, so its E_l is taken as 0 for the rest of the run (mlmc_input_t's
zero_protection field, mlmc.hpp:238, is read nowhere).  For the indicator
phi a level difference is nonzero only when the fine and coarse inner
estimates fall on opposite sides of the threshold, which at level 3 happens
for about one outer sample in 2000.  With the capped 3-sigma rule (cap 2^6),
N0 = 32, beta = 2, tau = 0.5, max_loss = 2 and seed 7, the driver's first
level-3 call (M0 = 1024 outer samples, ids 0..1023) draws no such sample: the
run keeps E_3 = 0 where the level's mean is -2.48e-5, which accounts for most
of the 3.5e-5 gap between the estimate and Phi(-2)."""
import json
import os

import numpy as np

from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_first_level3_batch_of_the_driver_has_no_differences():
    """The oracle's level-3 call of the driver's first M0 = 1024 outer samples
    (batched keying, ids from 0): all 1024 fine/coarse differences are zero
    (diffsums[0] = diffsums[1] = 0), while the fine indicators are not (their
    mean is 26/1024).  The next 16384 outer samples of the same level show
    nonzero differences: the level's mean is not zero, the first batch just
    drew none of them."""
    O = Oracle("f64").new(1, 32, 2, 0.0, True, moments=4, adapt_rule=1, cap_log2=6)
    O.model_init(1, 0, 0.5, 2.0, 7)
    first = O.sample(3, 1024, keying=1, outer_offset=0)
    assert first["diffsums"][0] == 0.0 and first["diffsums"][1] == 0.0
    assert first["fsums"][0] == 26.0
    later = O.sample(3, 16384, keying=1, outer_offset=1024)
    assert later["diffsums"][1] > 0.0
    assert later["diffsums"][0] < 0.0  # E_3 < 0, as the quadrature value -2.48e-5


def test_reference_driver_never_resamples_a_zero_variance_level(nm):
    """The reference's mlmc_t::run on a deterministic sampler whose level 3 has
    E_3 = V_3 = 0 (golden call sequence from the reference build): level 3 is
    sampled once, with M0, while its neighbours are sampled many times; the
    library's driver makes the same calls."""
    with open(os.path.join(ROOT, "tests", "golden", "reference_f64.json")) as f:
        fx = {c["case"]["name"]: c for c in json.load(f)["mlmc"]}
    case = fx["zero-variance-level"]
    cs = case["case"]
    l3 = [int(M) for ell, M in zip(case["ell"], case["M"]) if ell == 3]
    assert l3 == [cs["M0"]]
    for nb in (2, 4):
        assert sum(int(M) for ell, M in zip(case["ell"], case["M"]) if ell == nb) > 100 * cs["M0"]

    from test_capi import _two_point
    calls = []
    L = len(cs["E"])

    def level(ell, M):
        l = min(ell, L - 1)
        calls.append((ell, M))
        return dict(sums_fine=_two_point(cs["Ef"][l], cs["Vf"][l], M),
                    sums_diff=_two_point(cs["E"][l], cs["V"][l], M),
                    work=cs["W"][l] * M, M=M)

    r = nm.mlmc_run_cb(level, cs["tol"], cs["M0"])
    assert calls == [(e, int(M)) for e, M in zip(case["ell"], case["M"])]
    assert r["levels"][3]["M"] == cs["M0"]
    assert np.isclose(r["stat_error"], case["summary"][3], rtol=1e-15)
