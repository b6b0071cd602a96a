"""ctypes bindings for the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark, never by the product package.  Each wrapper
names the reference function it restates (paths relative to the fastem tree).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64


class MT64(ctypes.Structure):
    """std::mt19937_64 state (Util/Rand.cpp:7-14)."""
    _fields_ = [("mt", ctypes.c_uint64 * 312), ("mti", ctypes.c_int)]


def build():
    """Compile liboracle.so from oracle.c with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        mp = ctypes.POINTER(MT64)
        sig = {
            "orc_mt64_seed": (None, [mp, ctypes.c_uint64]),
            "orc_mt64_next": (ctypes.c_uint64, [mp]),
            "orc_mt64_canonical": (ctypes.c_double, [mp]),
            "orc_find_items": (None, [_i64, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
            "orc_parse_block_params": (ctypes.c_int, [ctypes.c_char_p, _i64, ctypes.c_int,
                                                      ctypes.c_int, _ip, _ip]),
            "orc_factor": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
            "orc_sym_eigvals": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
            "orc_cond": (ctypes.c_double, [ctypes.c_int, _dp]),
            "orc_regularize": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]),
            "orc_estep": (ctypes.c_double, [_i64, ctypes.c_int, ctypes.c_int, _fp, _dp, _dp,
                                            _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]),
            "orc_mstep": (None, [_i64, ctypes.c_int, ctypes.c_int, _fp, _dp, _i64, _dp, _dp,
                                 _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_train": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int, _fp, _dp, _dp, _dp,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_classify": (None, [_i64, ctypes.c_int, ctypes.c_int, _fp, _dp, _ip]),
            "orc_kmeanspp": (None, [_i64, ctypes.c_int, ctypes.c_int, _fp, mp, _dp]),
            "orc_lloyd": (ctypes.c_int, [_i64, ctypes.c_int, ctypes.c_int, _fp, _dp,
                                         ctypes.c_double, ctypes.c_int, mp, _dp]),
            "orc_gmm_from_clusters": (None, [_i64, ctypes.c_int, ctypes.c_int, _fp, _dp, _ip,
                                             _dp, _dp, _dp]),
            "orc_forgy": (None, [_i64, ctypes.c_int, ctypes.c_int, _fp, mp, _dp]),
            "orc_data_trace": (ctypes.c_double, [_i64, ctypes.c_int, _fp]),
            "orc_jain": (None, [_i64, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_double, mp, _dp,
                                _dp, _dp]),
            "orc_kmeansoror": (_i64, [_i64, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_double, mp,
                                      _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Rand:
    """The reference's global Rand() stream (Util/Rand.cpp:7-14,33-36)."""

    def __init__(self, seed=5489):
        self.state = MT64()
        lib().orc_mt64_seed(ctypes.byref(self.state), seed)

    def next_u64(self):
        return lib().orc_mt64_next(ctypes.byref(self.state))

    def canonical(self):
        return lib().orc_mt64_canonical(ctypes.byref(self.state))


def find_items(n, size, ident):
    """Util/OpenMP.h:37-46 FindOpenMPItems."""
    o, c = _i64(), _i64()
    lib().orc_find_items(n, size, ident, ctypes.byref(o), ctypes.byref(c))
    return o.value, c.value


def parse_block_params(spec, n, d, K):
    """EM2/BlockManager.cpp:18-28; raises ValueError like std::invalid_argument."""
    a, b = ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_parse_block_params(None if spec is None else spec.encode(), n, d, K,
                                      ctypes.byref(a), ctypes.byref(b))
    if rc != 0:
        raise ValueError("Bad value of MStepLoopParam")
    return a.value, b.value


def factor(C):
    """EM2/NormalDensity.cpp:43-65 -> (U = L^-T, logdet); raises on non-PD."""
    C = _c64(C)
    d = C.shape[0]
    U = np.zeros((d, d))
    ld = ctypes.c_double()
    if lib().orc_factor(d, _d(C), _d(U), ctypes.byref(ld)) != 0:
        raise np.linalg.LinAlgError("Cholesky decomposition failed")
    return U, ld.value


def factor_all(covs):
    K, d, _ = covs.shape
    U = np.zeros((K, d, d))
    ld = np.zeros(K)
    for k in range(K):
        U[k], ld[k] = factor(covs[k])
    return U, ld


def sym_eigvals(C):
    C = _c64(C)
    d = C.shape[0]
    ev = np.zeros(d)
    if lib().orc_sym_eigvals(d, _d(C), _d(ev)) != 0:
        raise RuntimeError("QL iteration did not converge")
    return ev


def cond(C):
    """EM2/NormalDensity.cpp:67-71."""
    C = _c64(C)
    return lib().orc_cond(C.shape[0], _d(C))


def regularize(covs, g):
    """EM2/EMAlgorithm.cpp:237-244 (returns a new array)."""
    c = _c64(covs).copy()
    K, d, _ = c.shape
    lib().orc_regularize(K, d, g, _d(c))
    return c


def estep(X, w, mu, U, logdet, nvt=16, nblk=1):
    """EM2/MatrixEM.cpp:39-87,275-307 -> (LL, R)."""
    X = _c32(X)
    n, d = X.shape
    K = len(w)
    R = np.zeros((n, K))
    w, mu, U, logdet = _c64(w), _c64(mu), _c64(U), _c64(logdet)
    ll = lib().orc_estep(n, d, K, _f(X), _d(w), _d(mu), _d(U), _d(logdet), _d(R), nvt, nblk)
    return ll, R


def mstep(X, R, ntotal=None, nvt=16, nblk=1, log2red=False):
    """EM2/MatrixEM.cpp:309-378 -> (w, mu, cov)."""
    X = _c32(X)
    R = _c64(R)
    n, d = X.shape
    K = R.shape[1]
    w = np.zeros(K)
    mu = np.zeros((K, d))
    cov = np.zeros((K, d, d))
    lib().orc_mstep(n, d, K, _f(X), _d(R), n if ntotal is None else ntotal, _d(w), _d(mu),
                    _d(cov), nvt, nblk, 1 if log2red else 0)
    return w, mu, cov


class TrainResult:
    def __init__(self, rc, w, mu, cov, ll_trace, ll):
        self.rc = rc
        self.w, self.mu, self.cov = w, mu, cov
        self.ll_trace = ll_trace
        self.ll = ll

    @property
    def iter_counter(self):
        return self.rc


def train(X, w, mu, cov, maxsweeps, eps=1e-5, athr=1e5, g=0.0, svd=True, nvt=16, nblk=1,
          log2red=False):
    """EM2/EMAlgorithm.cpp:133-191.  rc >= 0 is IterCounter; -1 Cholesky
    failure; -2 condition-number abort."""
    X = _c32(X)
    n, d = X.shape
    w, mu, cov = _c64(w).copy(), _c64(mu).copy(), _c64(cov).copy()
    K = len(w)
    tr = np.full(max(maxsweeps, 1), np.nan)
    ll = ctypes.c_double()
    rc = lib().orc_train(n, d, K, _f(X), _d(w), _d(mu), _d(cov), maxsweeps, eps, athr, g,
                         1 if svd else 0, nvt, nblk, 1 if log2red else 0, _d(tr),
                         ctypes.byref(ll))
    ntr = (rc - 1) if rc > 0 else 0
    return TrainResult(rc, w, mu, cov, tr[:ntr], ll.value)


def classify(X, C):
    X = _c32(X)
    C = _c64(C)
    n, d = X.shape
    cls = np.zeros(n, dtype=np.int32)
    lib().orc_classify(n, d, C.shape[0], _f(X), _d(C), cls.ctypes.data_as(_ip))
    return cls


def kmeanspp(X, K, rng):
    X = _c32(X)
    n, d = X.shape
    C = np.zeros((K, d))
    lib().orc_kmeanspp(n, d, K, _f(X), ctypes.byref(rng.state), _d(C))
    return C


def _seqdist(X64, c):
    """Squared distances of the rows to c, FP64, ascending feature order,
    no fused multiply-add (the device's and orc_kmeanspp's order)."""
    s = np.zeros(X64.shape[0])
    for j in range(X64.shape[1]):
        t = c[j] - X64[:, j]
        s = s + t * t
    return s


def kmeanspp_greedy(X, K, rng, L=None):
    """Greedy k-means++ (Arthur and Vassilvitskii 2007, section 6; no
    counterpart in the reference): the first seed as k-means++, then per seed
    L = 2 + floor(ln K) draws by the k-means++ D^2 rule (the first row i with
    r <= agg_i / agg_{n-1}, agg the sequential prefix sums), one variate
    each, and the candidate of least potential sum_i min(D_i^2, |x_i - c|^2)
    kept, ties to the earlier draw.  numpy restatement (test infrastructure)."""
    X64 = np.asarray(X, dtype=np.float32).astype(np.float64)
    n, d = X64.shape
    if L is None:
        L = min(8, 2 + int(np.log(K)))
    C = np.zeros((K, d))
    C[0] = X64[int(rng.canonical() * n)]
    dmin = None
    for k in range(1, K):
        dist = _seqdist(X64, C[k - 1])
        dmin = dist if dmin is None else np.where(dmin > dist, dist, dmin)
        agg = np.cumsum(dmin)
        S = agg[-1]
        cands = []
        for _ in range(L):
            r = rng.canonical()
            hit = np.nonzero(r <= agg / S)[0] if S > 0 else np.array([], dtype=np.int64)
            if hit.size:
                cands.append(int(hit[0]))
        if not cands:
            continue
        pots = [np.minimum(dmin, _seqdist(X64, X64[i])).sum() for i in cands]
        C[k] = X64[cands[int(np.argmin(pots))]]
    return C


def lloyd(X, C, rng, minrel=1e-4, maxiter=30):
    X = _c32(X)
    C = _c64(C).copy()
    n, d = X.shape
    bf = ctypes.c_double()
    it = lib().orc_lloyd(n, d, C.shape[0], _f(X), _d(C), minrel, maxiter,
                         ctypes.byref(rng.state), ctypes.byref(bf))
    return C, it, bf.value


def gmm_from_clusters(X, C, cls):
    X = _c32(X)
    C = _c64(C)
    n, d = X.shape
    K = C.shape[0]
    cls = np.ascontiguousarray(cls, dtype=np.int32)
    w = np.zeros(K)
    mu = np.zeros((K, d))
    cov = np.zeros((K, d, d))
    lib().orc_gmm_from_clusters(n, d, K, _f(X), _d(C), cls.ctypes.data_as(_ip), _d(w),
                                _d(mu), _d(cov))
    return w, mu, cov


def kmeanspp_init(X, K, seed):
    """k-means++ seeding, Lloyd (MinRel 1e-4, MaxIter 30), classify, and the
    mixture tail of EM2/EMInitializer.cpp:103-134, on one Rand() stream."""
    rng = Rand(seed)
    C = kmeanspp(X, K, rng)
    C, _, _ = lloyd(X, C, rng)
    cls = classify(X, C)
    return gmm_from_clusters(X, C, cls)


def forgy(X, K, rng):
    """Clust/KMeansInitializer.cpp:33-49: K rows at (long)(Rand()*N)."""
    X = _c32(X)
    n, d = X.shape
    C = np.zeros((K, d))
    lib().orc_forgy(n, d, K, _f(X), ctypes.byref(rng.state), _d(C))
    return C


def data_trace(X):
    """Util/MultithreadedDataset.cpp:179-283 TraceCov."""
    X = _c32(X)
    n, d = X.shape
    return lib().orc_data_trace(n, d, _f(X))


def jain_init(X, K, rng, trace=None):
    """EM2/EMInitializer.cpp:43-80 -> (w, mu, cov)."""
    X = _c32(X)
    n, d = X.shape
    tr = data_trace(X) if trace is None else float(trace)
    w = np.zeros(K)
    mu = np.zeros((K, d))
    cov = np.zeros((K, d, d))
    lib().orc_jain(n, d, K, _f(X), tr, ctypes.byref(rng.state), _d(w), _d(mu), _d(cov))
    return w, mu, cov


def kmeansoror(X, K, rng, oversample=1.4):
    """Clust/KMeansOrOrInitializer.cpp:143-174 in one-thread order -> (C, m)
    with m the number of candidate centers."""
    X = _c32(X)
    n, d = X.shape
    C = np.zeros((K, d))
    m = lib().orc_kmeansoror(n, d, K, _f(X), oversample, ctypes.byref(rng.state), _d(C))
    if m < 0:
        raise MemoryError("orc_kmeansoror")
    return C, m


def _lloyd_mixture(X, C, rng):
    C, _, _ = lloyd(X, C, rng)
    cls = classify(X, C)
    return gmm_from_clusters(X, C, cls)


def kmeans_init(X, K, seed):
    """KMeansEMInitializer::Init (EM2/EMInitializer.cpp:82-134): Forgy seeding,
    Lloyd (MinRel 1e-4, MaxIter 30), classify, cluster mixture; one stream."""
    rng = Rand(seed)
    return _lloyd_mixture(X, forgy(X, K, rng), rng)


def kmeansoror_init(X, K, seed):
    """KMeansOrOrEMInitializer (EM2/EMInitializer.cpp:158-166): k-means||
    (oversampling 1.4), then the KMeansEMInitializer tail."""
    rng = Rand(seed)
    C, _ = kmeansoror(X, K, rng)
    return _lloyd_mixture(X, C, rng)
