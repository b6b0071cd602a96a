"""fem_estep_noalloc (GPU): the E-step over caller-owned device buffers with
all working memory in the library's static device array.  Checked against
the context path (fem_factor + fem_estep + fem_get_posteriors) on the same
data and parameters: posteriors bit for bit, log-likelihood within 1e-12
relative (only the summation order over row tiles differs).  Two calls in a
row, the second smaller, show that nothing stale from the first call leaks
into the second."""
import ctypes

import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    fem.lib()  # loads the HIP runtime libfastem.so is linked against
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


class Dev:
    """A caller-owned device copy of a host array (the test allocates; the
    entry point under test does not)."""

    def __init__(self, h, a):
        self.h, self.shape, self.dtype = h, a.shape, a.dtype
        self.nbytes = a.nbytes
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), a.nbytes) == 0
        assert h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def context_estep(X, w, mu, cov):
    n, d = X.shape
    with fem.Context(n, d, len(w)) as ctx:
        ctx.upload(X)
        ctx.set_params(w, mu, cov)
        ctx.factor()
        ll = ctx.estep()
        R = ctx.posteriors()
    return ll, R


def noalloc_estep(h, X, w, mu, cov, ldx, ldr):
    n, d = X.shape
    K = len(w)
    Xp = np.zeros((n, ldx), np.float32)
    Xp[:, :d] = X
    bufs = [Dev(h, Xp)] + [Dev(h, np.ascontiguousarray(a, np.float64)) for a in (w, mu, cov)]
    bufs.append(Dev(h, np.full((n, ldr), -7.0)))  # sentinel in the pad columns
    ll = ctypes.c_double()
    info = np.full(K, -1, np.int32)
    rc = fem.lib().fem_estep_noalloc(bufs[0].p, n, ldx, d, K, bufs[1].p, bufs[2].p, bufs[3].p,
                                     bufs[4].p, ldr, ctypes.byref(ll),
                                     info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None)
    R = bufs[4].get()
    for b in bufs:
        b.free()
    return rc, ll.value, R, info


def test_noalloc_matches_context_twice():
    h = hip()
    cases = [(200000, 64, 64, 8, 64, 64, 4), (7000, 5, 3, 9, 8, 4, 1)]  # n, d, K, seed, ldx, ldr, tiles
    for n, d, K, seed, ldx, ldr, tiles in cases:
        print(f"noalloc n={n} d={d} K={K} ldx={ldx} ldr={ldr}", flush=True)
        X, w, mu, cov = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, seed)
        ll_ref, R_ref = context_estep(X, w, mu, cov)
        rc, ll, R, info = noalloc_estep(h, X, w, mu, cov, ldx, ldr)
        assert rc == fem.FEM_OK, fem.lib().fem_last_error(None)
        assert np.all(info == 0)
        np.testing.assert_array_equal(R[:, :K], R_ref)
        assert np.all(R[:, K:] == -7.0)  # pad columns untouched
        assert abs(ll - ll_ref) <= 1e-12 * abs(ll_ref), (ll, ll_ref)
        assert (n + 65535) // 65536 == tiles


def test_noalloc_errors():
    h = hip()
    X = np.zeros((100, 4), np.float32)
    w, mu, cov = np.full(2, 0.5), np.zeros((2, 4)), np.stack([np.eye(4)] * 2)
    rc = fem.lib().fem_estep_noalloc(None, 100, 4, 129, 2, None, None, None, None, 2, None, None,
                                     None)
    assert rc == fem.FEM_ERR_ARG
    cov[1, 3, 3] = -1.0
    rc, _, R, info = noalloc_estep(h, X, w, mu, cov, 4, 2)
    assert rc == fem.FEM_ERR_NOT_PD
    assert list(info) == [0, 4]
    assert b"component 1" in fem.lib().fem_last_error(None)
    assert np.all(R == -7.0)  # nothing written
