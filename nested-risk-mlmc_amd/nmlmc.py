"""Python bindings of libnmlmc.so, the MI355X nested sampler.

The C ABI (include/nmlmc.h) mirrors the reference's extern "C" entry points
(common.cpp:20-31, model_sampler.cpp:94-114, portfolio_sampler.cpp:618-756)
with its -DREAL_DOUBLE layouts; this module exposes them with the same names
plus small helpers.  The library is the only compute path: if it is missing,
loading fails loudly (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NMLMC_LIBRARY selects another build of the library (e.g. one compiled with other
# tuning macros); by default the in-tree build.
LIB_PATH = os.environ.get("NMLMC_LIBRARY") or os.path.join(HERE, "lib", "libnmlmc.so")

FORM_PLAIN, FORM_BATCHED, FORM_MIXED = 0, 1, 2
PHI_INDICATOR, PHI_MAX, PHI_MOMENTS, PHI_FMOMENTS = 0, 1, 2, 3
MODEL_REF, MODEL_SHIFT = 0, 1
METHOD_BS, METHOD_EXACT, METHOD_UNBIASED = 0, 1, 2
CALL, PUT, DIGITAL_CALL, DIGITAL_PUT = 1, -1, 2, -2


class NestedSamplerParams(C.Structure):
    """nested_sampler_params (nested_sampler.hpp:14-34), REAL_DOUBLE layout."""
    _fields_ = [("N0", C.c_uint32), ("beta", C.c_uint32), ("antithetic", C.c_bool),
                ("fine_moments", C.c_uint32), ("diff_moments", C.c_uint32),
                ("adaptive_r", C.c_double)]


class SampleOutput(C.Structure):
    """sample_output (common.hpp:37-55)."""
    _fields_ = [("M", C.c_uint64), ("fsums", C.POINTER(C.c_double)),
                ("diffsums", C.POINTER(C.c_double)), ("total_work", C.c_double),
                ("inner_stats", C.c_uint64 * 6)]


class ModelData(C.Structure):
    """model::data (model_sampler.cpp:21-39)."""
    _fields_ = [("stau", C.c_double), ("s1tau", C.c_double), ("max_loss", C.c_double),
                ("phi_type", C.c_int32)]


class Span(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t)]


class OptionParams(C.Structure):
    """portfolio::option_params (portfolio_sampler.cpp:42-52), REAL_DOUBLE layout."""
    _fields_ = [("strike", C.c_double), ("maturity", C.c_double), ("start", C.c_double),
                ("w", C.c_double), ("delta_est", C.c_double), ("stock_ind", C.c_uint32),
                ("type", C.c_int32), ("prob", C.c_double), ("W_est", C.c_double)]


class StockParams(C.Structure):
    _fields_ = [("S0", C.c_double), ("mu", C.c_double), ("sigma", C.c_double)]


class PortfolioData(C.Structure):
    """portfolio::data (portfolio_sampler.cpp:62-92), REAL_DOUBLE layout."""
    _fields_ = [("options", Span * 3), ("stocks", Span), ("r", C.c_double), ("rho", C.c_double),
                ("s1_rho2", C.c_double), ("max_loss", C.c_double), ("tau", C.c_double),
                ("dt_h0inv", C.c_uint32), ("dt_beta", C.c_uint32), ("dt_prob_expo", C.c_double),
                ("phi_type", C.c_int32), ("_phi_size", C.c_uint32), ("subsample", C.c_int32),
                ("use_delta_cv", C.c_bool), ("use_coarse_delta_cv", C.c_bool),
                ("use_antithetic_risk", C.c_bool), ("use_milstein", C.c_bool)]


class LevelStats(C.Structure):
    """nmlmc_level_stats: mlmc_t::mc_diff_output_t (mlmc.hpp:21-27), 4 moments."""
    _fields_ = [("sums_fine", C.c_double * 4), ("sums_diff", C.c_double * 4),
                ("work", C.c_double), ("time", C.c_double), ("M", C.c_uint64)]


SAMPLE_LEVEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint64, C.POINTER(LevelStats))

EXPORTS = [
    "init_nestedmc_sampler", "cleanup_nestedmc_sampler", "model_update_data", "model_init",
    "model_sample", "portfolio_cleanup", "portfolio_update_data", "portfolio_init",
    "portfolio_comp_params", "portfolio_sample", "nmlmc_last_error", "nmlmc_last_error_string",
    "nmlmc_set_form", "nmlmc_get_form", "nmlmc_set_adaptive_rule", "nmlmc_model_init",
    "nmlmc_portfolio_set_correlation", "nmlmc_sample", "nmlmc_mixed_stats",
    "nmlmc_last_kernel_ms", "nmlmc_last_kernel_name", "nmlmc_last_group_lanes",
    "nmlmc_mlmc_run", "nmlmc_mlmc_run_cb", "portfolio_sample_delta_unbiased",
    "portfolio_estimate_delta_for_unbiased", "portfolio_gen_data", "portfolio_free_data",
    "nmlmc_mlmc_last_table", "nmlmc_autotune_level", "nmlmc_sample_levels",
    "nmlmc_batch_begin", "nmlmc_batch_stage", "nmlmc_batch_sample", "nmlmc_batch_mlmc_run",
    "nmlmc_batch_autotune_level",
    "nmlmc_source_digest", "nmlmc_sample_noalloc", "nmlmc_sampler_model_init", "nmlmc_sampler_model_sample",
    "nmlmc_set_level_coupled_paths",
]

_lib = None


class NmlmcError(RuntimeError):
    pass


def source_digest():
    """sha256 (first 16 hex digits) of the library's sources as the Makefile
    computes it: csrc/*.hip, *.cpp, *.hpp and include/nmlmc.h, *.hpp,
    concatenated in path order; None when the sources are not beside the
    package."""
    import glob
    import hashlib
    files = [p for ext in ("hip", "cpp", "hpp") for p in glob.glob(os.path.join(HERE, "csrc", "*." + ext))]
    inc = os.path.join(os.path.dirname(HERE), "include")
    files.append(os.path.join(inc, "nmlmc.h"))
    files += glob.glob(os.path.join(inc, "*.hpp"))
    files = sorted(os.path.abspath(p) for p in files)
    if not all(os.path.exists(p) for p in files):
        return None
    h = hashlib.sha256()
    for p in files:
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def load(path=LIB_PATH):
    """Load libnmlmc.so (raises if it has not been built, or if the in-tree
    build does not match the sources beside it)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NmlmcError("%s is missing: build it with `make -C nested-risk-mlmc_amd`" % path)
    lib = C.CDLL(path)
    lib.nmlmc_source_digest.restype = C.c_char_p
    built = lib.nmlmc_source_digest().decode()
    want = source_digest()
    if want is not None and built != want and not os.environ.get("NMLMC_LIBRARY"):
        raise NmlmcError("%s was built from other sources (digest %s, sources %s): rebuild it with "
                         "`make -C nested-risk-mlmc_amd`" % (path, built, want))
    u32, u64, dbl, i32 = C.c_uint32, C.c_uint64, C.c_double, C.c_int
    dp, u64p = C.POINTER(C.c_double), C.POINTER(C.c_uint64)
    sig = {
        "init_nestedmc_sampler": (None, [u32, C.POINTER(NestedSamplerParams)]),
        "cleanup_nestedmc_sampler": (None, []),
        "model_update_data": (None, [C.POINTER(ModelData)]),
        "model_init": (None, [C.POINTER(ModelData), u64]),
        "model_sample": (None, [u32, u64, C.POINTER(SampleOutput)]),
        "portfolio_cleanup": (None, []),
        "portfolio_update_data": (None, [C.POINTER(PortfolioData)]),
        "portfolio_init": (None, [C.POINTER(PortfolioData), u64]),
        "portfolio_comp_params": (None, [C.POINTER(PortfolioData)]),
        "portfolio_sample": (None, [u32, u64, u64, C.POINTER(SampleOutput)]),
        "portfolio_sample_delta_unbiased": (None, [u32, u64p, dp, u32, dp]),
        "portfolio_estimate_delta_for_unbiased": (None, [C.POINTER(PortfolioData), dbl]),
        "portfolio_gen_data": (PortfolioData, [u32, u32, u64, dp, C.c_bool, C.c_bool]),
        "portfolio_free_data": (None, [C.POINTER(PortfolioData)]),
        "nmlmc_mlmc_last_table": (C.c_char_p, []),
        "nmlmc_last_error": (i32, []),
        "nmlmc_last_error_string": (C.c_char_p, []),
        "nmlmc_set_form": (i32, [i32]),
        "nmlmc_get_form": (i32, []),
        "nmlmc_set_adaptive_rule": (i32, [u32, u32]),
        "nmlmc_model_init": (i32, [i32, i32, dbl, dbl, u64]),
        "nmlmc_portfolio_set_correlation": (i32, [dp, u32]),
        "nmlmc_sample": (i32, [u32, u64, u64, C.POINTER(SampleOutput), dp, dp, u64p, u64p]),
        "nmlmc_mixed_stats": (i32, [u64p, u64p, u64p]),
        "nmlmc_last_kernel_ms": (dbl, []),
        "nmlmc_last_kernel_name": (C.c_char_p, []),
        "nmlmc_source_digest": (C.c_char_p, []),
        "nmlmc_sample_noalloc": (i32, [u32, u64, C.POINTER(SampleOutput)]),
        "nmlmc_set_level_coupled_paths": (i32, [i32]),
        "nmlmc_sampler_model_init": (i32, [C.POINTER(ModelData), u64]),
        "nmlmc_sampler_model_sample": (i32, [u32, u64, C.POINTER(SampleOutput)]),
        "nmlmc_last_group_lanes": (u32, []),
        "nmlmc_autotune_level": (i32, [u32, u64, C.POINTER(u32), C.POINTER(u32), dp]),
        "nmlmc_sample_levels": (i32, [u32, C.POINTER(u32), C.POINTER(u64), C.POINTER(SampleOutput)]),
        "nmlmc_batch_begin": (i32, [u32]),
        "nmlmc_batch_stage": (i32, [u32, C.POINTER(PortfolioData), u64, dp, u32]),
        "nmlmc_batch_sample": (i32, [u32, u32, C.POINTER(u64), C.POINTER(SampleOutput)]),
        "nmlmc_batch_autotune_level": (i32, [u32, u32, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32), dp]),
        "nmlmc_batch_mlmc_run": (i32, [u32, dp, u64, dbl, dbl, dbl, u32, C.POINTER(LevelStats),
                                       C.POINTER(u32), dp]),
        "nmlmc_mlmc_run": (i32, [dbl, u64, dbl, dbl, dbl, i32, u32, C.POINTER(LevelStats),
                                 C.POINTER(C.c_uint32), dp]),
        "nmlmc_mlmc_run_cb": (i32, [SAMPLE_LEVEL_FN, C.c_void_p, dbl, u64, dbl, dbl, dbl, i32, u32,
                                    C.POINTER(LevelStats), C.POINTER(C.c_uint32), dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc=None):
    """Raise NmlmcError if the last call failed."""
    lib = load()
    code = lib.nmlmc_last_error() if rc is None else rc
    if code:
        raise NmlmcError("nmlmc error %d: %s" % (code, lib.nmlmc_last_error_string().decode()))


class Output:
    """Owns the arrays of a sample_output."""

    def __init__(self, moments=4, phi_size=1):
        n = moments * phi_size
        self._f = (C.c_double * n)()
        self._d = (C.c_double * n)()
        self.c = SampleOutput()
        self.c.fsums = C.cast(self._f, C.POINTER(C.c_double))
        self.c.diffsums = C.cast(self._d, C.POINTER(C.c_double))
        self.fsums = np.frombuffer(self._f, np.float64)
        self.diffsums = np.frombuffer(self._d, np.float64)

    def as_dict(self):
        return dict(M=self.c.M, fsums=self.fsums.copy(), diffsums=self.diffsums.copy(),
                    work=self.c.total_work,
                    stats=np.frombuffer(self.c.inner_stats, np.uint64).copy())


# ---- reference entry points ---------------------------------------------------

def init_nestedmc_sampler(cores, N0, beta, adaptive_r=0.0, antithetic=True, fine_moments=4,
                          diff_moments=4):
    lib = load()
    p = NestedSamplerParams(N0, beta, bool(antithetic), fine_moments, diff_moments, adaptive_r)
    lib.init_nestedmc_sampler(cores, C.byref(p))
    check()


def cleanup_nestedmc_sampler():
    load().cleanup_nestedmc_sampler()


def model_init(phi_type, tau, max_loss, seed):
    """model_init with model::data(phi_type, tau, max_loss) (model_sampler.cpp:24-29)."""
    d = ModelData(np.sqrt(tau), np.sqrt(1 - tau), max_loss, phi_type)
    load().model_init(C.byref(d), seed)
    check()


def model_sample(ell, M, moments=4):
    o = Output(moments)
    load().model_sample(ell, M, C.byref(o.c))
    check()
    return o.as_dict()


# ---- extensions ---------------------------------------------------------------

def set_form(form):
    check(load().nmlmc_set_form(form))


def set_adaptive_rule(rule, cap_log2=0):
    check(load().nmlmc_set_adaptive_rule(rule, cap_log2))


def model_init_ex(variant, phi_type, tau, max_loss, seed):
    check(load().nmlmc_model_init(variant, phi_type, tau, max_loss, seed))


def sample(ell, M, outer_offset=None, trace=False, moments=4, phi_size=1):
    """nmlmc_sample: one level call; returns sums, work, stats (and the trace)."""
    lib = load()
    o = Output(moments, phi_size)
    off = (1 << 64) - 1 if outer_offset is None else outer_offset
    tr = None
    ptrs = [None] * 4
    if trace:
        tr = (np.zeros(M), np.zeros(M), np.zeros(M, np.uint64), np.zeros(M, np.uint64))
        ptrs = [tr[0].ctypes.data_as(C.POINTER(C.c_double)),
                tr[1].ctypes.data_as(C.POINTER(C.c_double)),
                tr[2].ctypes.data_as(C.POINTER(C.c_uint64)),
                tr[3].ctypes.data_as(C.POINTER(C.c_uint64))]
    check(lib.nmlmc_sample(ell, M, off, C.byref(o.c), *ptrs))
    out = o.as_dict()
    if trace:
        out.update(phi_f=tr[0], phi_c=tr[1], nf=tr[2], nc=tr[3])
    return out


def sample_noalloc(ell, M, moments=4, phi_size=1):
    """nmlmc_sample_noalloc: a plain-form level call that allocates no device
    memory (its scratch is the library's static array)."""
    o = Output(moments, phi_size)
    check(load().nmlmc_sample_noalloc(ell, M, C.byref(o.c)))
    return o.as_dict()


def sampler_model_init(phi_type, tau, max_loss, seed):
    """model::sampler as a class of the reference's sampler concept, staged
    through nmlmc::init_samplers<SAMPLER> (nmlmc_sampler_model_init)."""
    d = ModelData(np.sqrt(tau), np.sqrt(1 - tau), max_loss, phi_type)
    check(load().nmlmc_sampler_model_init(C.byref(d), seed))


def sampler_model_sample(ell, M, moments=4):
    """One level call of that class through nmlmc::sample<SAMPLER>."""
    o = Output(moments)
    check(load().nmlmc_sampler_model_sample(ell, M, C.byref(o.c)))
    return o.as_dict()


def sample_levels(ells, Ms, moments=4, phi_size=1):
    """nmlmc_sample_levels: a level sweep (one captured graph in the batched
    forms); returns one dict per level call, as sample() does."""
    lib = load()
    n = len(ells)
    outs = [Output(moments, phi_size) for _ in range(n)]
    arr = (SampleOutput * n)(*[o.c for o in outs])
    check(lib.nmlmc_sample_levels(n, (C.c_uint32 * n)(*ells), (C.c_uint64 * n)(*Ms), arr))
    res = []
    for i, o in enumerate(outs):
        o.c = arr[i]
        res.append(o.as_dict())
    return res


def mixed_stats():
    """Mixed form diagnostics of the last call: (outer samples re-evaluated in
    FP64, indicator flips, refinement decisions rechecked in FP64)."""
    a, b, r = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(load().nmlmc_mixed_stats(C.byref(a), C.byref(b), C.byref(r)))
    return a.value, b.value, r.value


def last_kernel_ms():
    return load().nmlmc_last_kernel_ms()


def last_kernel_name():
    return load().nmlmc_last_kernel_name().decode()


def last_group_lanes():
    return load().nmlmc_last_group_lanes()


def autotune_level(ell, M):
    """Time the batched level call (ell, M) over launch shapes and keep the
    fastest for later calls at this level: returns (segment, blocks_per_cu, ms)."""
    S, B, ms = C.c_uint32(), C.c_uint32(), C.c_double()
    check(load().nmlmc_autotune_level(ell, M, C.byref(S), C.byref(B), C.byref(ms)))
    return S.value, B.value, ms.value


def _run_outputs(lib_call, max_levels):
    levels = (LevelStats * max_levels)()
    n = C.c_uint32()
    summ = (C.c_double * 5)()
    rc = lib_call(levels, n, summ)
    check(rc)
    lv = []
    for i in range(min(n.value, max_levels)):
        L = levels[i]
        lv.append(dict(sums_fine=list(L.sums_fine), sums_diff=list(L.sums_diff), work=L.work,
                       time=L.time, M=L.M))
    return dict(levels=lv, starting_level=int(summ[0]), bias_est=summ[1], stat_error=summ[2],
                TOL=summ[3], wall=summ[4])


def mlmc_run(targetTOL, M0, theta=0.5, Ca=3.0, starting_level_coeff=1.5, verbose=False,
             max_levels=64):
    """mlmc_t<double,4>::run over the staged problem (mlmc.hpp:290-337)."""
    lib = load()
    return _run_outputs(lambda lv, n, s: lib.nmlmc_mlmc_run(
        targetTOL, M0, theta, Ca, starting_level_coeff, int(verbose), max_levels, lv,
        C.byref(n), s), max_levels)


def mlmc_run_cb(fn, targetTOL, M0, theta=0.5, Ca=3.0, starting_level_coeff=1.5, verbose=False,
                max_levels=64):
    """The same driver over a Python level sampler fn(ell, M) -> dict(sums_fine, sums_diff,
    work, M[, time])."""
    lib = load()

    def _cb(user, ell, M, out):
        try:
            r = fn(ell, M)
            for k in range(4):
                out.contents.sums_fine[k] = r["sums_fine"][k]
                out.contents.sums_diff[k] = r["sums_diff"][k]
            out.contents.work = r["work"]
            out.contents.time = r.get("time", 0.0)
            out.contents.M = r["M"]
            return 0
        except Exception:  # noqa: BLE001 - reported as an aborted run
            return 1

    cb = SAMPLE_LEVEL_FN(_cb)
    return _run_outputs(lambda lv, n, s: lib.nmlmc_mlmc_run_cb(
        cb, None, targetTOL, M0, theta, Ca, starting_level_coeff, int(verbose), max_levels, lv,
        C.byref(n), s), max_levels)


# ---- portfolio ----------------------------------------------------------------

OPTION_DTYPE = np.dtype([("strike", "<f8"), ("maturity", "<f8"), ("start", "<f8"), ("w", "<f8"),
                         ("delta_est", "<f8"), ("stock_ind", "<u4"), ("type", "<i4"),
                         ("prob", "<f8"), ("W_est", "<f8")])
STOCK_DTYPE = np.dtype([("S0", "<f8"), ("mu", "<f8"), ("sigma", "<f8")])


class Portfolio:
    """Owns the arrays behind a portfolio_data (portfolio_sampler.cpp:62-92):
    option records in a numpy array with the option_params layout (64 B).

    Options are given flat, method-major (BS, Exact, Unbiased) with ``counts``
    per method, each method's options sorted by descending ``prob``."""

    SCALARS = dict(r=0.05, rho=0.2, max_loss=0.0, tau=0.02, dt_h0inv=8, dt_beta=4,
                   dt_prob_expo=-1.5, phi_type=0, phi_size=1, subsample=1, use_delta_cv=True,
                   use_coarse_delta_cv=False, use_antithetic_risk=True, use_milstein=True)

    def __init__(self, S0, mu, sigma, counts, strike, maturity, w, stock_ind, type, prob,
                 W_est=None, start=None, delta_est=None, **scalars):
        assert OPTION_DTYPE.itemsize == C.sizeof(OptionParams)
        P = int(np.sum(counts))
        self.counts = [int(c) for c in counts]
        self.stocks = np.zeros(len(S0), STOCK_DTYPE)
        self.stocks["S0"], self.stocks["mu"], self.stocks["sigma"] = S0, mu, sigma
        self.ops = np.zeros(max(P, 1), OPTION_DTYPE)
        o = self.ops[:P]
        o["strike"], o["maturity"], o["w"] = strike, maturity, w
        o["stock_ind"], o["type"], o["prob"] = stock_ind, type, prob
        o["W_est"] = 1.0 if W_est is None else W_est
        o["start"] = 0.0 if start is None else start
        o["delta_est"] = 0.0 if delta_est is None else delta_est
        sc = dict(self.SCALARS)
        sc.update(scalars)
        self.scalars = sc
        self.c = PortfolioData()
        base = self.ops.ctypes.data
        off = 0
        for m in range(3):
            self.c.options[m] = Span(base + off * OPTION_DTYPE.itemsize, self.counts[m])
            off += self.counts[m]
        self.c.stocks = Span(self.stocks.ctypes.data, len(S0))
        self.c.r, self.c.rho = sc["r"], sc["rho"]
        self.c.s1_rho2 = float(np.sqrt(1 - sc["rho"] ** 2))
        self.c.max_loss, self.c.tau = sc["max_loss"], sc["tau"]
        self.c.dt_h0inv, self.c.dt_beta, self.c.dt_prob_expo = (sc["dt_h0inv"], sc["dt_beta"],
                                                                 sc["dt_prob_expo"])
        self.c.phi_type, self.c._phi_size, self.c.subsample = (sc["phi_type"], sc["phi_size"],
                                                               sc["subsample"])
        self.c.use_delta_cv = bool(sc["use_delta_cv"])
        self.c.use_coarse_delta_cv = bool(sc["use_coarse_delta_cv"])
        self.c.use_antithetic_risk = bool(sc["use_antithetic_risk"])
        self.c.use_milstein = bool(sc["use_milstein"])

    @property
    def phi_size(self):
        return 1 if self.scalars["phi_type"] in (0, 1) else self.scalars["phi_size"]

    @property
    def P(self):
        return sum(self.counts)

    def column(self, name):
        return self.ops[name][:self.P].copy()

    def set_max_loss(self, max_loss):
        self.scalars["max_loss"] = max_loss
        self.c.max_loss = max_loss

    def comp_params(self):
        """portfolio_comp_params: start = BS value, delta_est = BS delta."""
        load().portfolio_comp_params(C.byref(self.c))
        check()
        return self.column("start"), self.column("delta_est")


def portfolio_init(pf, seed):
    load().portfolio_init(C.byref(pf.c), seed)
    check()


def portfolio_update_data(pf):
    load().portfolio_update_data(C.byref(pf.c))
    check()


def portfolio_sample(ell, M, override_dyn_inner=0, moments=4, phi_size=1):
    o = Output(moments, phi_size)
    load().portfolio_sample(ell, override_dyn_inner, M, C.byref(o.c))
    check()
    return o.as_dict()


def set_level_coupled_paths(on):
    """Level-coupled inner paths for the method-2 options of portfolios staged
    from now on (nmlmc_set_level_coupled_paths)."""
    check(load().nmlmc_set_level_coupled_paths(int(bool(on))))


def portfolio_set_correlation(chol):
    lib = load()
    if chol is None:
        check(lib.nmlmc_portfolio_set_correlation(None, 0))
        return
    a = np.ascontiguousarray(chol, np.float64)
    check(lib.nmlmc_portfolio_set_correlation(a.ctypes.data_as(C.POINTER(C.c_double)), a.shape[0]))


def portfolio_cleanup():
    load().portfolio_cleanup()


def batch_begin(n):
    """Open a batch of n portfolio problems (nmlmc_batch_begin)."""
    check(load().nmlmc_batch_begin(n))


def batch_stage(i, pf, seed, chol=None):
    """Stage instance i of the batch: tables of Portfolio pf, its seed and
    optional Cholesky factor."""
    lib = load()
    if chol is None:
        check(lib.nmlmc_batch_stage(i, C.byref(pf.c), seed, None, 0))
        return
    a = np.ascontiguousarray(chol, np.float64)
    check(lib.nmlmc_batch_stage(i, C.byref(pf.c), seed, a.ctypes.data_as(C.POINTER(C.c_double)), a.shape[0]))


def batch_mlmc_run(targetTOLs, M0, theta=0.5, Ca=3.0, starting_level_coeff=1.5, max_levels=64):
    """The MLMC drivers of the staged batch run together (nmlmc_batch_mlmc_run):
    one result dict per instance (as mlmc_run gives), wall = the whole batch's."""
    lib = load()
    n = len(targetTOLs)
    levels = (LevelStats * (n * max_levels))()
    nl = (C.c_uint32 * n)()
    summ = (C.c_double * (5 * n))()
    check(lib.nmlmc_batch_mlmc_run(n, (C.c_double * n)(*targetTOLs), M0, theta, Ca, starting_level_coeff,
                                   max_levels, levels, nl, summ))
    res = []
    for i in range(n):
        lv = []
        for l in range(min(nl[i], max_levels)):
            L = levels[i * max_levels + l]
            lv.append(dict(sums_fine=list(L.sums_fine), sums_diff=list(L.sums_diff), work=L.work,
                           time=L.time, M=L.M))
        s = summ[5 * i:5 * i + 5]
        res.append(dict(levels=lv, starting_level=int(s[0]), bias_est=s[1], stat_error=s[2], TOL=s[3],
                        wall=s[4]))
    return res


def mlmc_estimate(r):
    """The MLMC estimate of a driver result: the starting level's fine mean plus
    the higher levels' difference means."""
    s = r["starting_level"]
    e = r["levels"][s]["sums_fine"][0] / r["levels"][s]["M"]
    for lv in r["levels"][s + 1:]:
        e += lv["sums_diff"][0] / lv["M"]
    return e


def batch_autotune_level(ell, Ms):
    """nmlmc_batch_autotune_level: time the batch level call (ell, Ms) over launch
    shapes and keep the fastest for later batch calls at this level; returns
    (segment, blocks_per_cu, ms)."""
    n = len(Ms)
    S, B, ms = C.c_uint32(), C.c_uint32(), C.c_double()
    check(load().nmlmc_batch_autotune_level(ell, n, (C.c_uint64 * n)(*Ms), C.byref(S), C.byref(B), C.byref(ms)))
    return S.value, B.value, ms.value


def batch_sample(ell, Ms, moments=4, phi_size=1):
    """One level call over every instance of the batch: one dict per instance."""
    lib = load()
    n = len(Ms)
    outs = [Output(moments, phi_size) for _ in range(n)]
    arr = (SampleOutput * n)(*[o.c for o in outs])
    check(lib.nmlmc_batch_sample(ell, n, (C.c_uint64 * n)(*Ms), arr))
    res = []
    for i, o in enumerate(outs):
        o.c = arr[i]
        res.append(o.as_dict())
    return res


def portfolio_sample_delta_unbiased(option_index, M, moments=2):
    """Moments of M delta samples of Unbiased option ``option_index`` and the work."""
    sums = np.zeros(moments)
    work = C.c_double(0)
    m = C.c_uint64(M)
    load().portfolio_sample_delta_unbiased(option_index, C.byref(m), sums.ctypes.data_as(
        C.POINTER(C.c_double)), moments, C.byref(work))
    check()
    return sums, work.value


def portfolio_estimate_delta_for_unbiased(pf, TOL):
    """Adds the estimated pathwise delta to each Unbiased option of ``pf``."""
    load().portfolio_estimate_delta_for_unbiased(C.byref(pf.c), TOL)
    check()


def _flatten_portfolio_data(pd):
    """Arrays of a portfolio_data in staged order (method-major)."""
    Q = pd.stocks.size
    st = C.cast(pd.stocks.data, C.POINTER(StockParams)) if Q else None
    out = {"S0": np.array([st[i].S0 for i in range(Q)]), "mu": np.array([st[i].mu for i in range(Q)]),
           "sigma": np.array([st[i].sigma for i in range(Q)]),
           "counts": np.array([pd.options[m].size for m in range(3)], np.uint32)}
    cols = {k: [] for k in ("strike", "maturity", "start", "w", "delta_est", "stock_ind", "type", "prob",
                            "W_est")}
    for m in range(3):
        n = pd.options[m].size
        if not n:
            continue
        ops = C.cast(pd.options[m].data, C.POINTER(OptionParams))
        for j in range(n):
            for k in cols:
                cols[k].append(getattr(ops[j], k))
    for k, v in cols.items():
        out[k] = np.array(v, dtype=np.uint32 if k == "stock_ind" else np.int32 if k == "type" else np.float64)
    scal = {k: getattr(pd, k) for k in ("r", "rho", "s1_rho2", "max_loss", "tau", "dt_h0inv", "dt_beta",
                                         "dt_prob_expo", "phi_type", "subsample", "use_delta_cv",
                                         "use_coarse_delta_cv", "use_antithetic_risk", "use_milstein")}
    scal["phi_size"] = pd._phi_size
    return out, scal


def portfolio_gen_data(P, Q, seed, probs, uniform_weights=False, uniform_prob=False):
    """portfolio_gen_data: returns (arrays, scalars) and frees the C arrays."""
    lib = load()
    pr = np.ascontiguousarray(probs, np.float64)
    pd = lib.portfolio_gen_data(P, Q, seed, pr.ctypes.data_as(C.POINTER(C.c_double)), uniform_weights,
                                uniform_prob)
    check()
    try:
        return _flatten_portfolio_data(pd)
    finally:
        lib.portfolio_free_data(C.byref(pd))


def mlmc_last_table():
    """mlmc_output_t::to_string of the last completed MLMC run (the level table)."""
    return load().nmlmc_mlmc_last_table().decode()
