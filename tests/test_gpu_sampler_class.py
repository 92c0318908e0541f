"""The reference's sampler concept on the batched engine: model::sampler
(model_sampler.cpp:42-92) written as a sampler class against
nmlmc::inner_sampler_base and run through nmlmc::init_samplers<SAMPLER> /
nmlmc::sample<SAMPLER> (include/nmlmc_sampler.hpp), against the oracle's
reference model under the batched keying and against the library's own
batched FP64 form of that model.  Counts, work and draw statistics exact; FP64
level sums to 1e-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CALLS = [(0, 300), (1, 200), (3, 60), (1, 50), (0, 7)]


def _cmp(g, o):
    assert np.array_equal(g["stats"], o["stats"]), (g["stats"], o["stats"])
    assert g["work"] == o["work"]
    np.testing.assert_allclose(g["fsums"], o["fsums"], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g["diffsums"], o["diffsums"], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("anti,phi,r", [(True, 0, 1.5), (True, 1, 1.5), (False, 0, 1.5), (True, 0, 0.0)])
def test_model_sampler_class_matches_oracle(nm, oracle_f64, anti, phi, r):
    """Level calls with continuing outer ids per level, antithetic and plain
    fine/coarse estimators, Indicator and Max, adaptive and not."""
    print("case", anti, phi, r, flush=True)
    tau, seed = 0.02, 29
    orc = oracle_f64.new(1, 32, 4, r, anti)
    ml = tau * (orc.norm_cdf_inv(0.025 / 2) ** 2 - 1)
    orc.model_init(0, phi, tau, ml, seed)
    nm.init_nestedmc_sampler(1, 32, 4, r, anti)
    nm.sampler_model_init(phi, tau, ml, seed)
    offsets = {}
    for ell, M in CALLS:
        off = offsets.get(ell, 0)
        g = nm.sampler_model_sample(ell, M)
        o = orc.sample(ell, M, keying=1, outer_offset=off)
        offsets[ell] = off + M
        _cmp(g, o)
    nm.cleanup_nestedmc_sampler()


def test_model_sampler_class_equals_batched_form(nm):
    """The class through the concept and the library's own batched FP64 form of
    the reference model draw the same samples: the same results."""
    tau, seed, ml = 0.02, 41, 0.09
    nm.init_nestedmc_sampler(1, 32, 4, 1.5, True)
    nm.sampler_model_init(0, tau, ml, seed)
    got = [nm.sampler_model_sample(ell, M) for ell, M in CALLS]
    nm.cleanup_nestedmc_sampler()
    nm.init_nestedmc_sampler(1, 32, 4, 1.5, True)
    nm.set_form(nm.FORM_BATCHED)
    nm.model_init(0, tau, ml, seed)
    want = [nm.sample(ell, M) for ell, M in CALLS]
    nm.cleanup_nestedmc_sampler()
    nm.set_form(nm.FORM_PLAIN)
    for g, w in zip(got, want):
        _cmp(g, w)


def test_sampler_class_staging_errors(nm):
    """A staged class is sampled through its template entry only; without
    staging, the class entry refuses."""
    nm.init_nestedmc_sampler(1, 32, 4, 1.5, True)
    with pytest.raises(nm.NmlmcError):
        nm.sampler_model_sample(0, 10)
    nm.sampler_model_init(0, 0.02, 0.1, 3)
    with pytest.raises(nm.NmlmcError):
        nm.sample(0, 10)
    nm.model_init(0, 0.02, 0.1, 3)  # a problem staged again: the class is no longer
    with pytest.raises(nm.NmlmcError):
        nm.sampler_model_sample(0, 10)
    nm.cleanup_nestedmc_sampler()
