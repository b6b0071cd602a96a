"""Mixed-precision forms against the FP64 oracle (GPU).

FP32 form: FP32 MFMA contractions, FP32 posteriors, FP64 Cholesky / log-det /
cross-chunk reductions.  Stated tolerances (relative, max-norm) on one E-step +
M-step from the same parameters:
  log-likelihood <= 2e-6, weights <= 1e-5, means <= 1e-5, covariances <= 1e-4;
over a 20-iteration trajectory the per-iteration LL stays within 1e-5.
"""
import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("n,d,K,lo,hi,diag", [(100000, 16, 8, -10.0, 10.0, 0.1),
                                              (50000, 64, 16, -20.0, 20.0, 0.05),
                                              (20000, 17, 5, -10.0, 10.0, 0.1),
                                              (30000, 128, 4, -20.0, 20.0, 0.05)])
def test_fp32_one_step(orc, n, d, K, lo, hi, diag):
    print(f"fp32 one step n={n} d={d} K={K}", flush=True)
    X, w, mu, cov = fem.gengmm(n, d, K, lo, hi, diag, 77 + d)
    rng = np.random.default_rng(1)
    mu0 = mu + 0.3 * rng.standard_normal(mu.shape)
    cov0 = cov * 1.2
    w0 = np.full(K, 1.0 / K)
    U, ld = orc.factor_all(cov0)
    ll_ref, R_ref = orc.estep(X, w0, mu0, U, ld)
    w_ref, mu_ref, cov_ref = orc.mstep(X, R_ref)
    with fem.Context(n, d, K, fem.PREC_FP32) as ctx:
        ctx.upload(X)
        ctx.set_params(w0, mu0, cov0)
        ctx.factor()
        ll = ctx.estep()
        R = ctx.posteriors()
        w1, mu1, cov1 = ctx.mstep()
    errs = {"ll": abs(ll - ll_ref) / abs(ll_ref), "R": float(np.max(np.abs(R - R_ref))),
            "w": rel(w1, w_ref), "mu": rel(mu1, mu_ref), "cov": rel(cov1, cov_ref)}
    print("fp32 errors", errs, flush=True)
    assert errs["ll"] <= 2e-6
    assert errs["R"] <= 1e-3
    assert errs["w"] <= 1e-5 and errs["mu"] <= 1e-5 and errs["cov"] <= 1e-4


def test_fp32_trajectory(orc):
    n, d, K, iters = 100000, 16, 8, 20
    print("fp32 trajectory", flush=True)
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 11)[0]
    w0, mu0, cov0 = fem.kmeanspp_init(X, K, seed=5)
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, svd=False)
    trace = []
    with fem.Context(n, d, K, fem.PREC_FP32) as ctx:
        ctx.upload(X)
        fem.train(ctx, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, ll_trace=trace)
    err = np.abs(np.array(trace) - ref.ll_trace) / np.abs(ref.ll_trace)
    print("fp32 per-iteration LL rel err", err.max(), flush=True)
    assert err.max() <= 1e-5
