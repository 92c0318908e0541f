"""ctypes bindings for the oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, where the oracle is the checker.  Two libraries are exposed:

* ``Oracle``: the C restatement in oracle/nmlmc_oracle.c (liboracle_f64.so /
  liboracle_f32.so), built by oracle/Makefile;
* ``Ref``: the reference's own OpenMP path compiled in place into
  oracle/_ref/ (libref_f64.so / libref_f32.so), present only where the
  reference tree was available when it was built.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def build(ref=True):
    """Build the oracle (and, when the reference tree exists, oracle/_ref)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


class OrcParams(C.Structure):
    _fields_ = [("N0", C.c_uint32), ("beta", C.c_uint32), ("antithetic", C.c_int32),
                ("fine_moments", C.c_uint32), ("diff_moments", C.c_uint32),
                ("adaptive_r", C.c_double), ("adapt_rule", C.c_uint32),
                ("cap_log2", C.c_uint32)]


class _Common:
    """Building blocks shared by the oracle and the reference probes."""

    def __init__(self, lib, prefix):
        self.lib = lib
        p = prefix
        self._tf = getattr(lib, p + "threefry2x64")
        self._tf.argtypes = [_u64p, _u64p, _u64p]
        self._stream = getattr(lib, p + "stream")
        self._stream.argtypes = [C.c_uint64, C.c_uint64, _ip, C.c_int, _u64p, _dp]
        self._bm = getattr(lib, p + "boxmuller")
        self._bm.argtypes = [C.c_uint64, C.c_uint64, _dp]
        for name, res in (("u01_f64", C.c_double), ("uneg11_f64", C.c_double),
                          ("u01_f32", C.c_float)):
            f = getattr(lib, p + name)
            f.argtypes = [C.c_uint64]
            f.restype = res
            setattr(self, name, f)
        self._cdf = getattr(lib, p + "rand_int_cdf")
        self._cdf.argtypes = [C.c_double, _dp, C.c_uint32]
        self._cdf.restype = C.c_uint32
        self._pdf = getattr(lib, p + "rand_int_pdf")
        self._pdf.argtypes = [C.c_double, _dp, C.c_uint32]
        self._pdf.restype = C.c_uint64
        for name in ("norm_cdf_inv", "norm_cdf"):
            f = getattr(lib, p + name)
            f.argtypes = [C.c_double]
            f.restype = C.c_double
            setattr(self, name, f)
        self.uint_pow = getattr(lib, p + "uint_pow")
        self.uint_pow.argtypes = [C.c_uint64, C.c_uint64]
        self.uint_pow.restype = C.c_uint64
        for name in ("bs_price", "bs_delta"):
            f = getattr(lib, p + name)
            f.argtypes = [C.c_int] + [C.c_double] * 5
            f.restype = C.c_double
            setattr(self, name, f)
        self.simulate_stock = getattr(lib, p + "simulate_stock")
        self.simulate_stock.argtypes = [C.c_double] * 5
        self.simulate_stock.restype = C.c_double
        self._ssdt = getattr(lib, p + "simulate_stock_dt")
        self._ssdt.argtypes = [C.c_int, _dp, _dp, C.c_uint64, C.c_double, C.c_double,
                               C.c_double, C.c_uint32, C.c_int, C.c_uint32]
        self._ssdt.restype = C.c_double
        # The path averages belong to the Asian extension: the restatement has
        # them, the reference build does not.
        self._ssdta = getattr(lib, p + "simulate_stock_dt_avg", None)
        if self._ssdta is not None:
            self._ssdta.argtypes = [C.c_int, _dp, _dp, _dp, C.c_uint64, C.c_double, C.c_double,
                                    C.c_double, C.c_uint32, C.c_int, C.c_uint32]
            self._ssdta.restype = C.c_double

    def threefry(self, ctr, key):
        c = (C.c_uint64 * 2)(*ctr)
        k = (C.c_uint64 * 2)(*key)
        o = (C.c_uint64 * 2)()
        self._tf(c, k, o)
        return [o[0], o[1]]

    def stream(self, seed, core, kinds):
        kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        n = len(kinds)
        ou = np.zeros(n, np.uint64)
        orr = np.zeros(n, np.float64)
        self._stream(seed, core, _ptr(kinds, _ip), n, _ptr(ou, _u64p), _ptr(orr, _dp))
        return ou, orr

    def boxmuller(self, u0, u1):
        o = (C.c_double * 2)()
        self._bm(u0, u1, o)
        return o[0], o[1]

    def rand_int_cdf(self, u, cdf):
        cdf = np.ascontiguousarray(cdf, np.float64)
        return self._cdf(u, _ptr(cdf, _dp), len(cdf))

    def rand_int_pdf(self, u, pdf):
        pdf = np.ascontiguousarray(pdf, np.float64)
        return self._pdf(u, _ptr(pdf, _dp), len(pdf))

    def simulate_stock_dt(self, with_coarse, normals, S, N, mu, sig, t, size, milstein,
                          coarse_step):
        normals = np.ascontiguousarray(normals, np.float64)
        S = np.array(S, np.float64).copy()
        W = self._ssdt(int(with_coarse), _ptr(normals, _dp), _ptr(S, _dp), N, mu, sig, t,
                       size, int(milstein), coarse_step)
        return W, S

    def simulate_stock_dt_avg(self, with_coarse, normals, S, N, mu, sig, t, size, milstein,
                              coarse_step):
        """simulate_stock_dt with the paths' averages over their steps (the Asian
        extension): returns (W, S, A)."""
        normals = np.ascontiguousarray(normals, np.float64)
        S = np.array(S, np.float64).copy()
        A = np.zeros_like(S)
        W = self._ssdta(int(with_coarse), _ptr(normals, _dp), _ptr(S, _dp), _ptr(A, _dp), N, mu, sig,
                        t, size, int(milstein), coarse_step)
        return W, S, A


def _portfolio_arrays(pf):
    """Flatten a portfolio dict into contiguous arrays (method-major options)."""
    out = {}
    for k in ("S0", "mu", "sigma", "strike", "maturity", "start", "w", "delta_est", "prob",
              "W_est"):
        out[k] = np.ascontiguousarray(pf[k], np.float64)
    out["counts"] = np.ascontiguousarray(pf["counts"], np.uint32)
    out["stock_ind"] = np.ascontiguousarray(pf["stock_ind"], np.uint32)
    out["type"] = np.ascontiguousarray(pf["type"], np.int32)
    return out


class Oracle(_Common):
    """The C restatement."""

    def __init__(self, real="f64"):
        path = os.path.join(HERE, "liboracle_%s.so" % real)
        if not os.path.exists(path):
            build(ref=False)
        super().__init__(C.CDLL(path), "orc_")
        lib = self.lib
        lib.orc_new.argtypes = [C.c_uint32, C.POINTER(OrcParams)]
        lib.orc_new.restype = C.c_void_p
        lib.orc_free.argtypes = [C.c_void_p]
        lib.orc_model_init.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_uint64]
        lib.orc_portfolio_init.argtypes = (
            [C.c_void_p, C.c_uint32, _dp, _dp, _dp, _u32p, _dp, _dp, _dp, _dp, _dp, _u32p, _ip,
             _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_uint32,
             C.c_double, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
             C.c_uint64])
        lib.orc_sample.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, _dp,
                                   _dp, _dp, _u64p, _dp, _dp, _u64p, _u64p]
        lib.orc_samples_count.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
        lib.orc_sample_delta_unbiased.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32,
                                                  _dp, _dp]
        lib.orc_samples_count.restype = C.c_uint64
        lib.orc_outer_states.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64, _dp, _dp]
        lib.orc_set_coupled.argtypes = [C.c_void_p, C.c_int]
        lib.orc_real_bytes.restype = C.c_int
        self.real = real
        self.ctx = None

    def new(self, cores, N0, beta, adaptive_r=0.0, antithetic=True, moments=4, adapt_rule=0,
            cap_log2=0):
        if self.ctx:
            self.lib.orc_free(self.ctx)
        self.moments = moments
        self.params = OrcParams(N0, beta, int(antithetic), moments, moments, adaptive_r,
                                adapt_rule, cap_log2)
        self.ctx = self.lib.orc_new(cores, C.byref(self.params))
        self.phi_size = 1
        return self

    def model_init(self, variant, phi_type, tau, max_loss, seed):
        self.phi_size = 1
        self.lib.orc_model_init(self.ctx, variant, phi_type, tau, max_loss, seed)

    def portfolio_init(self, pf, seed, chol=None):
        a = _portfolio_arrays(pf)
        self._pf_keep = a
        ch = None if chol is None else np.ascontiguousarray(chol, np.float64)
        self._chol_keep = ch
        phi_type = pf.get("phi_type", 0)
        phi_size = pf.get("phi_size", 1)
        self.phi_size = 1 if phi_type in (0, 1) else phi_size
        self.lib.orc_portfolio_init(
            self.ctx, len(a["S0"]), _ptr(a["S0"], _dp), _ptr(a["mu"], _dp), _ptr(a["sigma"], _dp),
            _ptr(a["counts"], _u32p), _ptr(a["strike"], _dp), _ptr(a["maturity"], _dp),
            _ptr(a["start"], _dp), _ptr(a["w"], _dp), _ptr(a["delta_est"], _dp),
            _ptr(a["stock_ind"], _u32p), _ptr(a["type"], _ip), _ptr(a["prob"], _dp),
            _ptr(a["W_est"], _dp), pf["r"], pf["rho"], pf["max_loss"], pf["tau"],
            pf.get("dt_h0inv", 8), pf.get("dt_beta", 4), pf.get("dt_prob_expo", -1.5), phi_type,
            phi_size, pf.get("subsample", 1), int(pf.get("use_delta_cv", True)),
            int(pf.get("use_coarse_delta_cv", False)), int(pf.get("use_antithetic_risk", True)),
            int(pf.get("use_milstein", True)), _ptr(ch, _dp), seed)

    def sample(self, ell, M, keying=0, outer_offset=0, trace=False):
        nm = self.moments * self.phi_size
        fs = np.zeros(nm)
        ds = np.zeros(nm)
        work = np.zeros(1)
        stats = np.zeros(6, np.uint64)
        tr = None
        if trace:
            tr = (np.zeros(M), np.zeros(M), np.zeros(M, np.uint64), np.zeros(M, np.uint64))
        self.lib.orc_sample(self.ctx, keying, ell, M, outer_offset, _ptr(fs, _dp), _ptr(ds, _dp),
                            _ptr(work, _dp), _ptr(stats, _u64p),
                            *((_ptr(tr[0], _dp), _ptr(tr[1], _dp), _ptr(tr[2], _u64p),
                               _ptr(tr[3], _u64p)) if tr else (None,) * 4))
        out = dict(fsums=fs, diffsums=ds, work=work[0], stats=stats)
        if trace:
            out.update(phi_f=tr[0], phi_c=tr[1], nf=tr[2], nc=tr[3])
        return out

    def sample_delta_unbiased(self, option_index, M, moments=2):
        """portfolio_sample_delta_unbiased with the plain keying."""
        sums = np.zeros(moments)
        work = np.zeros(1)
        self.lib.orc_sample_delta_unbiased(self.ctx, option_index, C.c_uint64(M), moments,
                                           _ptr(sums, _dp), _ptr(work, _dp))
        return sums, work[0]

    def set_coupled(self, on):
        """Level-coupled paths for the staged portfolio's method-2 options."""
        self.lib.orc_set_coupled(self.ctx, int(bool(on)))

    def outer_states(self, ell, M, outer_offset=0):
        """Portfolio outer draws under the batched keying: (R_tau [M x Q], the
        sampler's constant [M])."""
        Q = len(self._pf_keep["S0"])
        R = np.zeros((M, Q))
        const = np.zeros(M)
        self.lib.orc_outer_states(self.ctx, ell, M, outer_offset, _ptr(R, _dp), _ptr(const, _dp))
        return R, const

    def samples_count(self, ell):
        return (self.lib.orc_samples_count(self.ctx, ell, 0),
                self.lib.orc_samples_count(self.ctx, ell, 1))

    def __del__(self):
        try:
            if self.ctx:
                self.lib.orc_free(self.ctx)
        except Exception:
            pass


def ref_available(real="f64"):
    return os.path.exists(os.path.join(REF_DIR, "libref_%s.so" % real))


class Ref(_Common):
    """The reference's own OpenMP path, compiled in place (oracle/_ref)."""

    def __init__(self, real="f64"):
        path = os.path.join(REF_DIR, "libref_%s.so" % real)
        super().__init__(C.CDLL(path), "refp_")
        lib = self.lib
        lib.refp_set_threads.argtypes = [C.c_int]
        lib.refp_init_nested.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_int,
                                         C.c_uint32, C.c_uint32]
        lib.refp_model_init.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64]
        lib.refp_model_sample.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, _dp, _dp, _dp, _u64p]
        lib.refp_shift_init.argtypes = [C.c_double, C.c_uint64]
        lib.refp_shift_sample.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, _dp, _dp, _dp, _u64p]
        lib.refp_portfolio_set.argtypes = (
            [C.c_uint32, _dp, _dp, _dp, _u32p, _dp, _dp, _dp, _u32p, _ip, _dp, _dp, C.c_double,
             C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_int,
             C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int])
        lib.refp_portfolio_get_params.argtypes = [_dp, _dp]
        lib.refp_portfolio_init.argtypes = [C.c_uint64]
        lib.refp_portfolio_sample_delta.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, _dp, _dp]
        lib.refp_mlmc_last_table.restype = C.c_char_p
        lib.refp_portfolio_sample.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, _dp,
                                              _dp, _dp, _u64p]
        lib.refp_portfolio_gen.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, _dp, C.c_int,
                                           C.c_int, _dp, _dp, _dp, _u32p, _dp, _dp, _dp, _dp, _dp,
                                           _u32p, _ip, _dp, _dp]
        lib.refp_mlmc_fake_run.argtypes = [C.c_double, C.c_uint64, C.c_double, C.c_double,
                                           C.c_double, C.c_uint32, _dp, _dp, _dp, _dp, _dp,
                                           _u32p, _u64p, C.c_int, _dp]
        lib.refp_mlmc_fake_run.restype = C.c_int
        lib.refp_layout.argtypes = [_u64p]
        lib.refp_layout.restype = C.c_int
        self.real = real

    def layout(self):
        o = np.zeros(64, np.uint64)
        n = self.lib.refp_layout(_ptr(o, _u64p))
        return [int(x) for x in o[:n]]

    def init(self, cores, N0, beta, adaptive_r=0.0, antithetic=True, moments=4):
        self.lib.refp_set_threads(cores)
        self.moments = moments
        self.phi_size = 1
        self.lib.refp_init_nested(cores, N0, beta, adaptive_r, int(antithetic), moments, moments)
        return self

    def model_init(self, phi_type, tau, max_loss, seed):
        self.phi_size = 1
        self.lib.refp_model_init(phi_type, tau, max_loss, seed)

    def model_sample(self, ell, M):
        fs = np.zeros(self.moments)
        ds = np.zeros(self.moments)
        work = np.zeros(1)
        stats = np.zeros(6, np.uint64)
        self.lib.refp_model_sample(ell, M, self.moments, _ptr(fs, _dp), _ptr(ds, _dp),
                                   _ptr(work, _dp), _ptr(stats, _u64p))
        return dict(fsums=fs, diffsums=ds, work=work[0], stats=stats)

    def shift_init(self, max_loss, seed):
        """The shift model X = Y + Z (Indicator phi) on the reference's engine."""
        self.phi_size = 1
        self.lib.refp_shift_init(max_loss, seed)

    def shift_sample(self, ell, M):
        fs = np.zeros(self.moments)
        ds = np.zeros(self.moments)
        work = np.zeros(1)
        stats = np.zeros(6, np.uint64)
        self.lib.refp_shift_sample(ell, M, self.moments, _ptr(fs, _dp), _ptr(ds, _dp), _ptr(work, _dp),
                                   _ptr(stats, _u64p))
        return dict(fsums=fs, diffsums=ds, work=work[0], stats=stats)

    def portfolio_set(self, pf):
        """Stage a portfolio; returns (start, delta_est) from portfolio_comp_params."""
        a = _portfolio_arrays(pf)
        self._pf_keep = a
        phi_type = pf.get("phi_type", 0)
        phi_size = pf.get("phi_size", 1)
        self.phi_size = 1 if phi_type in (0, 1) else phi_size
        self.lib.refp_portfolio_set(
            len(a["S0"]), _ptr(a["S0"], _dp), _ptr(a["mu"], _dp), _ptr(a["sigma"], _dp),
            _ptr(a["counts"], _u32p), _ptr(a["strike"], _dp), _ptr(a["maturity"], _dp),
            _ptr(a["w"], _dp), _ptr(a["stock_ind"], _u32p), _ptr(a["type"], _ip),
            _ptr(a["prob"], _dp), _ptr(a["W_est"], _dp), pf["r"], pf["rho"], pf["max_loss"],
            pf["tau"], pf.get("dt_h0inv", 8), pf.get("dt_beta", 4), pf.get("dt_prob_expo", -1.5),
            phi_type, phi_size, pf.get("subsample", 1), int(pf.get("use_delta_cv", True)),
            int(pf.get("use_coarse_delta_cv", False)), int(pf.get("use_antithetic_risk", True)),
            int(pf.get("use_milstein", True)))
        n = int(a["counts"].sum())
        start = np.zeros(n)
        delta = np.zeros(n)
        self.lib.refp_portfolio_get_params(_ptr(start, _dp), _ptr(delta, _dp))
        return start, delta

    def portfolio_init(self, seed):
        self.lib.refp_portfolio_init(seed)

    def portfolio_sample(self, ell, M):
        nm = self.moments * self.phi_size
        fs = np.zeros(nm)
        ds = np.zeros(nm)
        work = np.zeros(1)
        stats = np.zeros(6, np.uint64)
        self.lib.refp_portfolio_sample(ell, M, self.moments, self.phi_size, _ptr(fs, _dp),
                                       _ptr(ds, _dp), _ptr(work, _dp), _ptr(stats, _u64p))
        return dict(fsums=fs, diffsums=ds, work=work[0], stats=stats)

    def portfolio_sample_delta(self, option_index, M, moments=2):
        sums = np.zeros(moments)
        work = np.zeros(1)
        self.lib.refp_portfolio_sample_delta(option_index, C.c_uint64(M), moments, _ptr(sums, _dp),
                                             _ptr(work, _dp))
        return sums, work[0]

    def mlmc_last_table(self):
        return self.lib.refp_mlmc_last_table().decode()

    def portfolio_gen(self, P, Q, seed, probs, uniform_weights=False, uniform_prob=False):
        probs = np.ascontiguousarray(probs, np.float64)
        d = {k: np.zeros(Q) for k in ("S0", "mu", "sigma")}
        for k in ("strike", "maturity", "start", "w", "delta_est", "prob", "W_est"):
            d[k] = np.zeros(P)
        d["counts"] = np.zeros(3, np.uint32)
        d["stock_ind"] = np.zeros(P, np.uint32)
        d["type"] = np.zeros(P, np.int32)
        self.lib.refp_portfolio_gen(
            P, Q, seed, _ptr(probs, _dp), int(uniform_weights), int(uniform_prob),
            _ptr(d["S0"], _dp), _ptr(d["mu"], _dp), _ptr(d["sigma"], _dp), _ptr(d["counts"], _u32p),
            _ptr(d["strike"], _dp), _ptr(d["maturity"], _dp), _ptr(d["start"], _dp), _ptr(d["w"], _dp),
            _ptr(d["delta_est"], _dp), _ptr(d["stock_ind"], _u32p), _ptr(d["type"], _ip),
            _ptr(d["prob"], _dp), _ptr(d["W_est"], _dp))
        return d

    def mlmc_fake_run(self, targetTOL, M0, E, V, W, Ef, Vf, theta=0.5, Ca=3.0, slc=1.5,
                      max_calls=4096):
        arrs = [np.ascontiguousarray(x, np.float64) for x in (E, V, W, Ef, Vf)]
        ell = np.zeros(max_calls, np.uint32)
        M = np.zeros(max_calls, np.uint64)
        summ = np.zeros(5)
        n = self.lib.refp_mlmc_fake_run(targetTOL, M0, theta, Ca, slc, len(arrs[0]),
                                        *[_ptr(a, _dp) for a in arrs], _ptr(ell, _u32p),
                                        _ptr(M, _u64p), max_calls, _ptr(summ, _dp))
        n = min(n, max_calls)
        return ell[:n].copy(), M[:n].copy(), summ
