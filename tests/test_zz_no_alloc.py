"""The no-allocation entry point (nmlmc_sample_noalloc, no counterpart in the
reference) against the plain form's existing entry points, called through the
C ABI: model_sample and portfolio_sample of the same staged problem, from the
same seed, give the same level sums, work and counts, bit for bit (the same
kernels and the same fixed-order reduction), on two calls in a row, the second
with a smaller input; oversize grids and the batched forms are refused."""
import ctypes as C

import numpy as np
import pytest

from portfolio_util import synthetic, to_library

pytestmark = pytest.mark.gpu

CALLS = [(3, 70000), (1, 700)]  # the second call smaller; the first holds every virtual thread


def _run(nm, staging, cores, entry):
    nm.init_nestedmc_sampler(cores, 32, 4, 1.5, True)
    keep = staging()
    lib = nm.load()
    res = []
    for ell, M in CALLS:
        out = nm.Output(4, 1)
        if entry == "noalloc":
            rc = lib.nmlmc_sample_noalloc(ell, M, C.byref(out.c))
            assert rc == 0, lib.nmlmc_last_error_string()
        elif entry == "model":
            lib.model_sample(ell, M, C.byref(out.c))
            nm.check()
        else:
            lib.portfolio_sample(ell, 0, M, C.byref(out.c))
            nm.check()
        res.append(out.as_dict())
    del keep
    nm.cleanup_nestedmc_sampler()
    return res


def _same(a, b):
    assert a["M"] == b["M"]
    assert np.array_equal(a["stats"], b["stats"]), (a["stats"], b["stats"])
    assert a["work"] == b["work"]
    assert np.array_equal(a["fsums"], b["fsums"])
    assert np.array_equal(a["diffsums"], b["diffsums"])


@pytest.mark.parametrize("cores", [4096, 65536])
def test_noalloc_model_matches_model_sample(nm, cores):
    """Reference model, adaptive rule, 4096 and 65536 virtual threads (at level
    3 a thread takes 64 lanes: 65536 threads, all holding outer samples at
    M = 70000, fill the scratch's 16384 blocks)."""
    def stage():
        nm.model_init(0, 0.02, 0.02 * (2.2414 ** 2 - 1), 11)
        return None
    print("model, cores", cores, flush=True)
    want = _run(nm, stage, cores, "model")
    got = _run(nm, stage, cores, "noalloc")
    for a, b in zip(got, want):
        _same(a, b)


def test_noalloc_portfolio_matches_portfolio_sample(nm):
    """A sub-sampled BS + Exact book with digitals, 512 virtual threads."""
    d = synthetic(P=60, Q=3, counts=(10, 50, 0), seed=7)
    scal = dict(r=0.05, rho=0.2, tau=0.02, dt_h0inv=8, dt_beta=4, dt_prob_expo=-1.5, max_loss=0.01,
                subsample=1, use_delta_cv=True, use_antithetic_risk=True, use_coarse_delta_cv=False,
                use_milstein=True, phi_type=0, phi_size=1)

    def stage():
        pf, _, _ = to_library(nm, d, **scal)
        nm.portfolio_init(pf, 5)
        return pf
    print("portfolio", flush=True)
    want = _run(nm, stage, 512, "portfolio")
    got = _run(nm, stage, 512, "noalloc")
    for a, b in zip(got, want):
        _same(a, b)


def test_noalloc_refusals(nm):
    """More blocks than the scratch holds, a batched form: errors with a
    status, nothing launched.  Reference model, N0 = 32, beta = 4: a level-0
    thread takes 32 lanes (32 Box-Muller pairs), a level-1 thread 64.  Only the
    virtual threads holding outer samples (min(M, T)) are launched, so 131072
    of them need 16384 blocks at level 0 (the limit) and 32768 at level 1,
    while 100 need one or a few blocks at either level."""
    lib = nm.load()
    out = nm.Output(4, 1)
    nm.init_nestedmc_sampler(131072, 32, 4, 1.5, True)
    nm.model_init(0, 0.02, 0.1, 3)
    assert lib.nmlmc_sample_noalloc(0, 131072, C.byref(out.c)) == 0
    assert lib.nmlmc_sample_noalloc(1, 100, C.byref(out.c)) == 0
    assert lib.nmlmc_sample_noalloc(1, 131072, C.byref(out.c)) == 2  # NMLMC_ERR_ARG
    assert b"scratch holds" in lib.nmlmc_last_error_string()
    with pytest.raises(nm.NmlmcError):
        nm.sample_noalloc(1, 131072)
    nm.set_form(nm.FORM_BATCHED)
    assert lib.nmlmc_sample_noalloc(0, 100, C.byref(out.c)) == 1  # NMLMC_ERR_STATE
    nm.set_form(nm.FORM_PLAIN)
    nm.cleanup_nestedmc_sampler()
