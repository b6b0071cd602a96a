"""fastem_amd -- Gaussian-mixture EM engine for one AMD Instinct MI355X.

Python bindings (ctypes) of the C ABI in include/fastem_abi.h.  The compute
path is the HIP library libfastem.so built in-tree for gfx950; there is no CPU
fallback: if the library is missing, loading it raises.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libfastem.so")
EM2_PATH = os.path.join(_PKG, "em2")
EMINIT_PATH = os.path.join(_PKG, "eminit")

PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
ESTEP_FORMS = ("fp64", "fp32", "bf16x3", "bf16x3-ref")  # fem_estep_form
SEED_FORGY, SEED_KMEANSPP, SEED_KMEANSOROR, SEED_KMEANSPP_GREEDY = 0, 1, 2, 3
FEM_OK, FEM_ERR_ARG, FEM_ERR_HIP, FEM_ERR_NOT_PD, FEM_ERR_STATE = 0, -1, -2, -3, -4

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


class FemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fastem error {code}: {msg}")
        self.code = code


_SIGS = {
    "fem_create": (ctypes.c_int, [ctypes.POINTER(_vp), _i64, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int]),
    "fem_destroy": (ctypes.c_int, [_vp]),
    "fem_upload_data": (ctypes.c_int, [_vp, _fp, _i64]),
    "fem_upload_data_device": (ctypes.c_int, [_vp, _vp, _i64]),
    "fem_set_total_rows": (ctypes.c_int, [_vp, _i64]),
    "fem_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fem_set_regularization": (ctypes.c_int, [_vp, ctypes.c_double]),
    "fem_set_params": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "fem_get_params": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "fem_get_params_start": (ctypes.c_int, [_vp]),
    "fem_get_params_finish": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "fem_factor": (ctypes.c_int, [_vp, _dp, _ip]),
    "fem_estep": (ctypes.c_int, [_vp, _dp]),
    "fem_mstep": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "fem_get_posteriors": (ctypes.c_int, [_vp, _dp, _i64]),
    "fem_last_times": (ctypes.c_int, [_vp, _fp, _fp, _fp, _fp]),
    "fem_last_estep_form": (ctypes.c_int, [_vp, _ip]),
    "fem_last_estep_skipped": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "fem_kmeans_init": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                       _dp]),
    "fem_kmeans_init_rng": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                           _dp]),
    "fem_kmeans_init_seeded": (ctypes.c_int, [_vp, _dp, ctypes.c_uint64, ctypes.c_int, _dp, _dp,
                                              _dp]),
    "fem_last_error": (ctypes.c_char_p, [_vp]),
    "fem_get_classes": (ctypes.c_int, [_vp, _ip]),
    "fem_condition_number": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
    "fem_bhatt_dist": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
    "fem_estep_noalloc": (ctypes.c_int, [_vp, _i64, _i64, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                         _vp, _i64, _dp, _ip, _vp]),
    "fem_version": (ctypes.c_char_p, []),
    "fem_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int64), _ip, _ip, ctypes.c_int]),
    "fem_batch_destroy": (ctypes.c_int, [_vp]),
    "fem_batch_upload": (ctypes.c_int, [_vp, ctypes.c_int, _fp, _i64]),
    "fem_batch_set_params": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
    "fem_batch_tune": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int]),
    "fem_batch_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int]),
    "fem_batch_get_result": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip]),
    "fem_batch_last_error": (ctypes.c_char_p, [_vp]),
    "fem_gengmm": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _fp, _ip,
                                  _dp, _dp, _dp]),
    "fem_kmeanspp_init": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_uint64,
                                         ctypes.c_int, _dp, _dp, _dp]),
    "fem_setnet_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint64, ctypes.c_int]),
    "fem_setnet_destroy": (ctypes.c_int, [_vp]),
    "fem_setnet_last_error": (ctypes.c_char_p, [_vp]),
    "fem_setnet_num_params": (ctypes.c_int, [_vp]),
    "fem_setnet_get_params": (ctypes.c_int, [_vp, _fp]),
    "fem_setnet_set_params": (ctypes.c_int, [_vp, _fp]),
    "fem_setnet_forward": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, _fp]),
    "fem_setnet_forward_bf16": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, _fp]),
    "fem_setnet_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp,
                                       _fp, _fp, ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp,
                                       _fp]),
}


def lib():
    """Loads libfastem.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m fastem_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Context:
    """One device-resident dataset plus mixture parameters (fem_ctx)."""

    def __init__(self, n, d, K, precision=PREC_FP64, device=0):
        self.n, self.d, self.K = int(n), int(d), int(K)
        self.precision = precision
        h = _vp()
        rc = lib().fem_create(ctypes.byref(h), self.n, self.d, self.K, precision, device)
        if rc != 0:
            raise FemError(rc, lib().fem_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise FemError(rc, lib().fem_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().fem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        assert X.shape == (self.n, self.d)
        self._check(lib().fem_upload_data(self._h, X.ctypes.data_as(_fp), self.d))

    def upload_device(self, ptr, ld=None):
        self._check(lib().fem_upload_data_device(self._h, _vp(ptr), ld or self.d))

    def set_precision(self, precision):
        self._check(lib().fem_set_precision(self._h, precision))
        self.precision = precision

    def set_total_rows(self, ntotal):
        self._check(lib().fem_set_total_rows(self._h, int(ntotal)))

    def set_regularization(self, g):
        self._check(lib().fem_set_regularization(self._h, float(g)))

    def set_params(self, w, mu, cov):
        w = np.ascontiguousarray(w, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        cov = np.ascontiguousarray(cov, dtype=np.float64)
        assert w.shape == (self.K,) and mu.shape == (self.K, self.d)
        assert cov.shape == (self.K, self.d, self.d)
        self._check(lib().fem_set_params(self._h, _d(w), _d(mu), _d(cov)))

    def get_params(self):
        w = np.empty(self.K)
        mu = np.empty((self.K, self.d))
        cov = np.empty((self.K, self.d, self.d))
        self._check(lib().fem_get_params(self._h, _d(w), _d(mu), _d(cov)))
        return w, mu, cov

    def get_params_start(self):
        """Queues the parameter download into the context's pinned staging."""
        self._check(lib().fem_get_params_start(self._h))

    def get_params_finish(self):
        """Waits for the download get_params_start queued; (w, mu, cov)."""
        w = np.empty(self.K)
        mu = np.empty((self.K, self.d))
        cov = np.empty((self.K, self.d, self.d))
        self._check(lib().fem_get_params_finish(self._h, _d(w), _d(mu), _d(cov)))
        return w, mu, cov

    def factor(self):
        """Returns (logdet, info); raises FemError(FEM_ERR_NOT_PD) on failure
        after filling info."""
        ld = np.empty(self.K)
        info = np.zeros(self.K, dtype=np.int32)
        rc = lib().fem_factor(self._h, _d(ld), info.ctypes.data_as(_ip))
        if rc != 0:
            err = FemError(rc, lib().fem_last_error(self._h).decode())
            err.info = info
            raise err
        return ld, info

    def estep(self):
        ll = ctypes.c_double()
        self._check(lib().fem_estep(self._h, ctypes.byref(ll)))
        return ll.value

    def mstep(self, copy=True):
        if not copy:
            self._check(lib().fem_mstep(self._h, None, None, None))
            return None
        w = np.empty(self.K)
        mu = np.empty((self.K, self.d))
        cov = np.empty((self.K, self.d, self.d))
        self._check(lib().fem_mstep(self._h, _d(w), _d(mu), _d(cov)))
        return w, mu, cov

    def posteriors(self):
        R = np.empty((self.n, self.K))
        self._check(lib().fem_get_posteriors(self._h, _d(R), self.K))
        return R

    def kmeans_init(self, seed, seeding=1, max_sweeps=29):
        """Forgy (SEED_FORGY), k-means++ (SEED_KMEANSPP), k-means||
        (SEED_KMEANSOROR) or greedy k-means++ (SEED_KMEANSPP_GREEDY)
        seeding, Lloyd sweeps and the mixture of the clusters on the device
        data; becomes the parameters."""
        w = np.empty(self.K)
        mu = np.empty((self.K, self.d))
        cov = np.empty((self.K, self.d, self.d))
        self._check(lib().fem_kmeans_init(self._h, seed, seeding, max_sweeps, _d(w), _d(mu),
                                          _d(cov)))
        return w, mu, cov

    def kmeans_init_seeded(self, C0, seed=0, max_sweeps=29):
        """Lloyd sweeps and the cluster mixture from given centroids C0
        (K x d); becomes the parameters."""
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
        assert C0.shape == (self.K, self.d)
        w = np.empty(self.K)
        mu = np.empty((self.K, self.d))
        cov = np.empty((self.K, self.d, self.d))
        self._check(lib().fem_kmeans_init_seeded(self._h, _d(C0), seed, max_sweeps, _d(w), _d(mu),
                                                 _d(cov)))
        return w, mu, cov

    def classes(self):
        """Component of largest posterior per row (fem_get_classes)."""
        cls = np.empty(self.n, dtype=np.int32)
        self._check(lib().fem_get_classes(self._h, cls.ctypes.data_as(_ip)))
        return cls

    def times(self):
        t = [ctypes.c_float() for _ in range(4)]
        self._check(lib().fem_last_times(self._h, *[ctypes.byref(x) for x in t]))
        return {"estep_ms": t[0].value, "means_ms": t[1].value, "cov_ms": t[2].value,
                "factor_ms": t[3].value}

    def estep_skipped(self):
        """Share of the (wave, component) evaluations the last FP64 E-step
        skipped as provably beyond exp's underflow range (0 for other forms
        and for contexts with K > 64, d <= 16 or FASTEM_ESTEP_DENSE=1)."""
        f = ctypes.c_double()
        self._check(lib().fem_last_estep_skipped(self._h, ctypes.byref(f)))
        return f.value

    def estep_form(self):
        """Kernel of the last estep(): "fp64", "fp32", "bf16x3" (centred on
        each component) or "bf16x3-ref" (centred on a reference component)."""
        f = ctypes.c_int()
        self._check(lib().fem_last_estep_form(self._h, ctypes.byref(f)))
        return ESTEP_FORMS[f.value]


def condition_number(C):
    """|lambda_max / lambda_min| of a covariance, as em2's default check
    computes it (NormalDensity::ComputeConditionNumber); host only."""
    C = np.ascontiguousarray(C, dtype=np.float64)
    out = ctypes.c_double()
    rc = lib().fem_condition_number(C.shape[0], _d(C), ctypes.byref(out))
    if rc != 0:
        raise FemError(rc, "fem_condition_number failed")
    return out.value


def bhatt_dist(mu1, C1, mu2, C2):
    """Bhattacharyya distance of two normal densities (BhattDist); host only."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (mu1, C1, mu2, C2)]
    out = ctypes.c_double()
    rc = lib().fem_bhatt_dist(len(a[0]), *[_d(x) for x in a], ctypes.byref(out))
    if rc != 0:
        raise FemError(rc, "fem_bhatt_dist failed")
    return out.value


def regularize(cov, g):
    """EM2/EMAlgorithm.cpp:237-244 on host arrays (used before the first
    factorisation, as TrainNoCopy does at :228)."""
    cov = np.array(cov, dtype=np.float64, copy=True)
    d = cov.shape[-1]
    for k in range(cov.shape[0]):
        tr = np.trace(cov[k])
        cov[k] = cov[k] * (1.0 - g) + (g * tr / d) * np.eye(d)
    return cov


def train(ctx, w, mu, cov, maxsweeps, eps=1e-5, g=0.0, ll_trace=None):
    """EMAlgorithm::TrainNoCopy (EM2/EMAlgorithm.cpp:139-191) over the C ABI,
    without the condition-number check.  Returns (iter_counter, LL, (w, mu,
    cov)) with the reference's IterCounter semantics."""
    ctx.set_regularization(g)
    ctx.set_params(w, mu, regularize(cov, g))
    ctx.factor()
    prev = -np.finfo(np.float64).max
    ll = prev
    it = 1
    while it <= maxsweeps:
        ll = ctx.estep()
        if ll_trace is not None:
            ll_trace.append(ll)
        ctx.mstep(copy=False)
        ctx.factor()
        e = abs((ll - prev) / ll)
        if ll < prev:
            e = 0.0
        prev = ll
        if e < eps:
            it += 1
            break
        it += 1
    return it, ll, ctx.get_params()


def gengmm(n, d, K, mean_lo, mean_hi, diag, seed, labels=False):
    """Seeded synthetic mixture (include/fastem_tools.h).  Returns X (FP32)
    and the generating (w, mu, cov) [and labels]."""
    X = np.empty((n, d), dtype=np.float32)
    lab = np.empty(n, dtype=np.int32) if labels else None
    w = np.empty(K)
    mu = np.empty((K, d))
    cov = np.empty((K, d, d))
    rc = lib().fem_gengmm(n, d, K, mean_lo, mean_hi, diag, seed, X.ctypes.data_as(_fp),
                          None if lab is None else lab.ctypes.data_as(_ip), _d(w), _d(mu),
                          _d(cov))
    if rc != 0:
        raise FemError(rc, "fem_gengmm failed")
    return (X, w, mu, cov, lab) if labels else (X, w, mu, cov)


def kmeanspp_init(X, K, seed, lloyd_iters=29):
    """k-means++ seeding, up to lloyd_iters Lloyd sweeps, mixture from the
    clusters, on a temporary device context (include/fastem_tools.h)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    w = np.empty(K)
    mu = np.empty((K, d))
    cov = np.empty((K, d, d))
    rc = lib().fem_kmeanspp_init(n, d, K, X.ctypes.data_as(_fp), seed, lloyd_iters, _d(w),
                                 _d(mu), _d(cov))
    if rc != 0:
        raise FemError(rc, "k-means++ init left a cluster with fewer than two points")
    return w, mu, cov


class Batch:
    """Many independent fits of uneven (n, d <= 64, K) advanced together
    (fem_batch_*): one persistent launch per EM phase, one captured hipGraph
    per iteration of all fits, device-side convergence per fit."""

    CONVERGED, SWEEP_CAP, FAILED = 1, 2, -1

    def __init__(self, shapes, device=0):
        self.shapes = [(int(n), int(d), int(K)) for n, d, K in shapes]
        F = len(self.shapes)
        n = (ctypes.c_int64 * F)(*[s[0] for s in self.shapes])
        d = (ctypes.c_int * F)(*[s[1] for s in self.shapes])
        K = (ctypes.c_int * F)(*[s[2] for s in self.shapes])
        h = _vp()
        rc = lib().fem_batch_create(ctypes.byref(h), F, n, d, K, device)
        if rc != 0:
            raise FemError(rc, "fem_batch_create failed")
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise FemError(rc, lib().fem_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().fem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, fit, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, d, _ = self.shapes[fit]
        assert X.shape == (n, d)
        self._check(lib().fem_batch_upload(self._h, fit, X.ctypes.data_as(_fp), d))

    def set_params(self, fit, w, mu, cov):
        w = np.ascontiguousarray(w, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        cov = np.ascontiguousarray(cov, dtype=np.float64)
        self._check(lib().fem_batch_set_params(self._h, fit, _d(w), _d(mu), _d(cov)))

    def tune(self):
        buf = ctypes.create_string_buffer(65536)
        self._check(lib().fem_batch_tune(self._h, buf, 65536))
        return buf.value.decode()

    def run(self, maxsweeps, eps=1e-5, g=0.0, use_graph=True):
        self._check(lib().fem_batch_run(self._h, maxsweeps, eps, g, 1 if use_graph else 0))

    def result(self, fit):
        n, d, K = self.shapes[fit]
        w = np.empty(K)
        mu = np.empty((K, d))
        cov = np.empty((K, d, d))
        ll = ctypes.c_double()
        it = ctypes.c_int()
        st = ctypes.c_int()
        self._check(lib().fem_batch_get_result(self._h, fit, _d(w), _d(mu), _d(cov),
                                               ctypes.byref(ll), ctypes.byref(it),
                                               ctypes.byref(st)))
        return {"w": w, "mu": mu, "cov": cov, "ll": ll.value, "iter_counter": it.value,
                "status": st.value}


def bic(ll, n, d, K):
    """Bayesian information criterion of a full-covariance mixture:
    -2 LL + p ln n with p = (K-1) + K d + K d (d+1)/2."""
    p = (K - 1) + K * d + K * d * (d + 1) // 2
    return -2.0 * ll + p * np.log(n)
