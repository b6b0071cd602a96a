"""Learned initializer: a permutation-invariant set network that proposes
starting means and diagonal covariances for EM (fem_setnet_* in
include/fastem_setnet.h; kernels in csrc/fem_setnet.hip).

The reference has no such initializer: it starts EM from the Jain, k-means,
k-means++ or file initializers (EM2/EMInitializer.cpp).  Here the network
proposes the starting means and they replace the k-means++ seeding pass; the
proposal itself runs in the library's C entries (fem_setnet_propose,
fem_setnet_judge; `init_context`, `judge`): Lloyd sweeps and the cluster
mixture follow on the device as for k-means++, k-means++ stays the fallback,
and `judge` keeps whichever of the learned and the k-means++ start reaches
the higher final EM log-likelihood.  em2 and eminit reach the same entries
through `-I learned:<file>` and `-I learnedbest:<file>`.

Two heads propose means: the seed-query head (K attention-pooled queries ->
means and log-variances) and the point head (a learned mean-shift step that
moves every point of the subsample towards its cluster's mean; k-means++
seeding over the moved points then picks one seed per cluster).  The point
head's proposal is the one the initializer uses.

Inference runs the network's products on BF16 MFMA (fem_setnet_forward_bf16);
training uses the FP32 forward.  Weights are stored as raw files
(fem_setnet_save / fem_setnet_load).

Coordinates: a point set X is normalised by its own centroid c and one
scale s (root mean square deviation over all coordinates), so the network
sees (X - c) / s and predicts means and log-variances in those units.
"""
import ctypes

import numpy as np

from . import FemError, _d, _fp, _ip, _vp, lib

_dp = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "fem_setnet_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "fem_setnet_load": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_char_p, ctypes.c_int]),
    "fem_setnet_shape": (ctypes.c_int, [_vp, _ip, _ip]),
    "fem_setnet_propose_means": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_uint64, _dp]),
    "fem_setnet_propose": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _dp, _dp,
                                          _dp, _ip]),
    "fem_setnet_judge": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _ip, _dp]),
}


def _lib():
    L = lib()
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def normalise(X):
    X = np.asarray(X, dtype=np.float64)
    c = X.mean(axis=0)
    s = float(np.sqrt(np.mean((X - c) ** 2)))
    if not s > 0:
        s = 1.0
    return (X - c) / s, c, s


def subsample(X, n, seed):
    """n rows of X drawn uniformly (with replacement if X has fewer)."""
    rng = np.random.default_rng(seed)
    N = X.shape[0]
    if N <= n:
        idx = rng.integers(0, N, size=n) if N < n else np.arange(N)
    else:
        idx = np.sort(rng.choice(N, size=n, replace=False))
    return X[idx]


def point_targets(Z, M):
    """Point-head targets: the nearest of the fitted means M (K x d) to every
    row of Z (same normalised units)."""
    D = (Z * Z).sum(1)[:, None] - 2.0 * Z @ M.T + (M * M).sum(1)[None, :]
    return D.argmin(axis=1)


def fit_to_targets(w, mu, cov, c, s):
    """A fitted mixture in the network's normalised units: (M, V, p)."""
    M = (np.asarray(mu) - c) / s
    V = np.log(np.maximum(np.diagonal(np.asarray(cov), axis1=1, axis2=2), 1e-300) / (s * s))
    return M, V, np.asarray(w, dtype=np.float64)


class SetNet:
    """FP32 set network on one HIP device (fem_setnet)."""

    def __init__(self, d, K, seed=0, device=0, _handle=None):
        self.d, self.K = int(d), int(K)
        if _handle is None:
            h = _vp()
            rc = lib().fem_setnet_create(ctypes.byref(h), self.d, self.K, int(seed), device)
            if rc != 0:
                raise FemError(rc, "fem_setnet_create failed")
            _handle = h
        self._h = _handle
        self.nparams = lib().fem_setnet_num_params(self._h)

    def _check(self, rc):
        if rc != 0:
            raise FemError(rc, lib().fem_setnet_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().fem_setnet_destroy(self._h)
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

    def params(self):
        th = np.empty(self.nparams, dtype=np.float32)
        self._check(lib().fem_setnet_get_params(self._h, th.ctypes.data_as(_fp)))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == self.nparams
        self._check(lib().fem_setnet_set_params(self._h, th.ctypes.data_as(_fp)))

    def save(self, path):
        """Raw weight file (fem_setnet_save)."""
        self._check(_lib().fem_setnet_save(self._h, str(path).encode()))

    @classmethod
    def load(cls, path, device=0):
        """Network from a raw weight file (fem_setnet_load)."""
        L = _lib()
        h = _vp()
        rc = L.fem_setnet_load(ctypes.byref(h), str(path).encode(), device)
        if rc != 0:
            raise FemError(rc, L.fem_last_error(None).decode())
        d, K = ctypes.c_int(), ctypes.c_int()
        L.fem_setnet_shape(h, ctypes.byref(d), ctypes.byref(K))
        return cls(d.value, K.value, _handle=h)

    def forward(self, Xs, points=False, bf16=False):
        """Xs: (B, n, d) normalised sets -> (mu, lv), each (B, K, d), and with
        points=True also the point head's outputs m (B, n, d).  bf16=True runs
        the inference forward with its products on BF16 MFMA
        (fem_setnet_forward_bf16); the default is the FP32 training forward."""
        Xs = np.ascontiguousarray(Xs, dtype=np.float32)
        B, n, d = Xs.shape
        assert d == self.d
        mu = np.empty((B, self.K, d), dtype=np.float32)
        lv = np.empty((B, self.K, d), dtype=np.float32)
        m = np.empty((B, n, d), dtype=np.float32) if points else None
        f = lib().fem_setnet_forward_bf16 if bf16 else lib().fem_setnet_forward
        self._check(f(self._h, B, n, Xs.ctypes.data_as(_fp), mu.ctypes.data_as(_fp),
                      lv.ctypes.data_as(_fp), m.ctypes.data_as(_fp) if points else None))
        return (mu, lv, m) if points else (mu, lv)

    def step(self, Xs, M, V, p, T=None, lam=0.5, gam=1.0, lr=0.0, want_grad=False):
        """Loss (summed over the batch) and optionally its gradient; with
        lr > 0 also one Adam update on the batch-mean loss.  T: point targets
        (B, n, d) or None (point head left out of the loss)."""
        Xs = np.ascontiguousarray(Xs, dtype=np.float32)
        M = np.ascontiguousarray(M, dtype=np.float32)
        V = np.ascontiguousarray(V, dtype=np.float32)
        p = np.ascontiguousarray(p, dtype=np.float32)
        B, n, d = Xs.shape
        J = M.shape[1]
        assert M.shape == (B, J, d) and V.shape == (B, J, d) and p.shape == (B, J)
        if T is not None:
            T = np.ascontiguousarray(T, dtype=np.float32)
            assert T.shape == (B, n, d)
        grad = np.empty(self.nparams, dtype=np.float32) if want_grad else None
        loss = ctypes.c_float()
        self._check(lib().fem_setnet_step(self._h, B, n, Xs.ctypes.data_as(_fp), J,
                                          M.ctypes.data_as(_fp), V.ctypes.data_as(_fp),
                                          p.ctypes.data_as(_fp),
                                          T.ctypes.data_as(_fp) if T is not None else None,
                                          float(lam), float(gam), float(lr),
                                          grad.ctypes.data_as(_fp) if want_grad else None,
                                          ctypes.byref(loss)))
        return (loss.value, grad) if want_grad else loss.value

    def propose_means(self, ctx, n_sub=2048, seed=0):
        """The network's K starting means (K x d, FP64) for the dataset held by
        ctx: the point head moves each point of an n_sub subsample towards its
        cluster's mean and k-means++ seeding over the moved points picks one
        per cluster (fem_setnet_propose_means; std::mt19937_64(seed) as the
        Rand() stream)."""
        C0 = np.empty((self.K, self.d))
        rc = _lib().fem_setnet_propose_means(self._h, ctx._h, int(n_sub), int(seed), _d(C0))
        if rc != 0:
            raise FemError(rc, lib().fem_last_error(ctx._h).decode())
        return C0

    def propose_query_means(self, X, n_sub=1024, seed=0, bf16=True):
        """The seed-query head's K means for dataset X (K x d, FP64)."""
        Xs = subsample(np.asarray(X), n_sub, seed)
        Z, c, s = normalise(Xs)
        mu, _ = self.forward(Z[None].astype(np.float32), bf16=bf16)
        return c + s * mu[0].astype(np.float64)

    def init_context(self, ctx, X=None, n_sub=2048, seed=0, lloyd_sweeps=29):
        """Learned initialisation of a device context (fem_setnet_propose):
        the proposed means seed Lloyd sweeps and the cluster mixture on the
        device in place of the k-means++ seeding pass; k-means++ on the same
        Rand() stream is the fallback if a cluster keeps fewer than two
        points, or, when the context has no covariance regularisation
        (Context.set_regularization), fewer than d + 1 (a singular
        covariance).  X (ignored) is accepted for older callers.  Returns
        (w, mu, cov, used) with used "learned" or "kmeans++"."""
        w, mu, cov = np.empty(self.K), np.empty((self.K, self.d)), np.empty((self.K, self.d, self.d))
        used = ctypes.c_int()
        rc = _lib().fem_setnet_propose(self._h, ctx._h, int(n_sub), int(seed), int(lloyd_sweeps), _d(w), _d(mu),
                                       _d(cov), ctypes.byref(used))
        if rc != 0:
            raise FemError(rc, lib().fem_last_error(ctx._h).decode())
        return w, mu, cov, ("learned" if used.value == 1 else "kmeans++")

    def judge(self, ctx, n_sub=2048, seed=0, lloyd_sweeps=29, em_sweeps=1000, eps=1e-5, g=0.0):
        """EM from the learned and from the k-means++ initialisation, the
        higher final log-likelihood kept (fem_setnet_judge).  Returns
        (w, mu, cov, winner, (ll_learned, ll_kmeanspp)) with the converged
        mixture; winner "learned" or "kmeans++"."""
        w, mu, cov = np.empty(self.K), np.empty((self.K, self.d)), np.empty((self.K, self.d, self.d))
        used = ctypes.c_int()
        ll = np.empty(2)
        rc = _lib().fem_setnet_judge(self._h, ctx._h, int(n_sub), int(seed), int(lloyd_sweeps), int(em_sweeps),
                                     float(eps), float(g), _d(w), _d(mu), _d(cov), ctypes.byref(used), _d(ll))
        if rc != 0:
            raise FemError(rc, lib().fem_last_error(ctx._h).decode())
        return w, mu, cov, ("learned" if used.value == 1 else "kmeans++"), (ll[0], ll[1])

    def propose(self, X, n_sub=1024, seed=0, var_floor=1e-6, bf16=True):
        """Starting mixture (w, mu, cov) for dataset X: uniform weights,
        network means, diagonal covariances exp(log-variance)."""
        Xs = subsample(np.asarray(X), n_sub, seed)
        Z, c, s = normalise(Xs)
        mu, lv = self.forward(Z[None].astype(np.float32), bf16=bf16)
        mu = c + s * mu[0].astype(np.float64)
        var = np.maximum(np.exp(np.clip(lv[0].astype(np.float64), -30, 30)) * s * s, var_floor)
        cov = np.zeros((self.K, self.d, self.d))
        idx = np.arange(self.d)
        cov[:, idx, idx] = var
        return np.full(self.K, 1.0 / self.K), mu, cov
