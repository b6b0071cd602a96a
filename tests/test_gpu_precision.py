"""Mixed-precision forms against the FP64 oracle (GPU).

FP32 form: FP32 MFMA contractions, FP32 posteriors, FP64 Cholesky / log-det /
cross-chunk reductions.  Stated tolerances (relative, max-norm) on one E-step +
M-step from the same parameters:
  log-likelihood <= 2e-6, weights <= 1e-5, means <= 1e-5, covariances <= 1e-4;
over a 20-iteration trajectory the per-iteration LL stays within 1e-5.

BF16x3 form: every contraction operand split into two bf16 (hi + lo), three
bf16 MFMAs per product with FP32 accumulation, FP32 posteriors, FP64 Cholesky
/ log-det / cross-chunk reductions.  Stated tolerances on one step:
  log-likelihood <= 2e-6, posteriors <= 5e-3 absolute, weights <= 1e-5,
  means <= 1e-5, covariances <= 2e-5;
per-iteration LL over a 20-iteration trajectory within 2e-6; at d = 128,
K = 128 the per-iteration errors are reported and bounded by 1e-4 (LL),
1e-4 (means) and 1e-4 (covariances).  The errors are printed.

The overlapping case (means in [-1.5, 1.5]^d, 0.5 I added to the
covariances) has fractional posteriors everywhere; in the separated cases
most posteriors are 0 or 1 in both forms.
"""
import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


TOL = {fem.PREC_FP32: {"ll": 2e-6, "R": 1e-3, "w": 1e-5, "mu": 1e-5, "cov": 1e-4, "traj": 1e-5},
       fem.PREC_BF16X3: {"ll": 2e-6, "R": 5e-3, "w": 1e-5, "mu": 1e-5, "cov": 2e-5, "traj": 2e-6}}
NAME = {fem.PREC_FP32: "fp32", fem.PREC_BF16X3: "bf16x3"}

SHAPES = [(100000, 16, 8, -10.0, 10.0, 0.1), (50000, 64, 16, -20.0, 20.0, 0.05),
          (20000, 17, 5, -10.0, 10.0, 0.1), (30000, 128, 4, -20.0, 20.0, 0.05),
          (50000, 32, 8, -1.5, 1.5, 0.5), (15000, 100, 3, -20.0, 20.0, 0.05),
          (40000, 8, 80, -20.0, 20.0, 0.05), (20000, 3, 4, -10.0, 10.0, 0.1),
          (20000, 1, 2, -10.0, 10.0, 0.1), (20000, 80, 5, -20.0, 20.0, 0.05),
          (15000, 96, 4, -20.0, 20.0, 0.05)]
BF16_SHAPES = [(40000, 48, 6, -10.0, 10.0, 0.1), (12000, 128, 16, -20.0, 20.0, 0.05)]


@pytest.mark.parametrize("prec,n,d,K,lo,hi,diag",
                         [(fem.PREC_FP32,) + s for s in SHAPES] +
                         [(fem.PREC_BF16X3,) + s for s in SHAPES + BF16_SHAPES])
def test_one_step(orc, prec, n, d, K, lo, hi, diag):
    print(f"{NAME[prec]} one step n={n} d={d} K={K}", flush=True)
    X, w, mu, cov = fem.gengmm(n, d, K, lo, hi, diag, 77 + d)
    rng = np.random.default_rng(1)
    mu0 = mu + 0.3 * rng.standard_normal(mu.shape)
    cov0 = cov * 1.2
    w0 = np.full(K, 1.0 / K)
    U, ld = orc.factor_all(cov0)
    ll_ref, R_ref = orc.estep(X, w0, mu0, U, ld)
    w_ref, mu_ref, cov_ref = orc.mstep(X, R_ref)
    with fem.Context(n, d, K, prec) as ctx:
        ctx.upload(X)
        ctx.set_params(w0, mu0, cov0)
        ctx.factor()
        ll = ctx.estep()
        R = ctx.posteriors()
        w1, mu1, cov1 = ctx.mstep()
    errs = {"ll": abs(ll - ll_ref) / abs(ll_ref), "R": float(np.max(np.abs(R - R_ref))),
            "w": rel(w1, w_ref), "mu": rel(mu1, mu_ref), "cov": rel(cov1, cov_ref)}
    print(NAME[prec], "errors", errs, flush=True)
    tol = TOL[prec]
    for key in errs:
        assert errs[key] <= tol[key], (key, errs[key])


@pytest.mark.parametrize("prec", [fem.PREC_FP32, fem.PREC_BF16X3])
def test_trajectory(orc, prec):
    n, d, K, iters = 100000, 16, 8, 20
    print(f"{NAME[prec]} trajectory", flush=True)
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 11)[0]
    w0, mu0, cov0 = fem.kmeanspp_init(X, K, seed=5)
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, svd=False)
    trace = []
    with fem.Context(n, d, K, prec) as ctx:
        ctx.upload(X)
        fem.train(ctx, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, ll_trace=trace)
    err = np.abs(np.array(trace) - ref.ll_trace) / np.abs(ref.ll_trace)
    print(NAME[prec], "per-iteration LL rel err", err.max(), flush=True)
    assert err.max() <= TOL[prec]["traj"]


def one_step(X, w, mu, cov, prec, dense):
    """Posteriors and M-step of one step; dense=True sums the covariances
    over every point (FASTEM_COV_DENSE=1), else over the nonzero posteriors."""
    import os
    n, d = X.shape
    if dense:
        os.environ["FASTEM_COV_DENSE"] = "1"
    try:
        with fem.Context(n, d, len(w), prec) as ctx:
            ctx.upload(X)
            ctx.set_params(w, mu, cov)
            ctx.factor()
            ctx.estep()
            return ctx.posteriors(), ctx.mstep()
    finally:
        os.environ.pop("FASTEM_COV_DENSE", None)


@pytest.mark.parametrize("prec", [fem.PREC_FP32, fem.PREC_BF16X3])
@pytest.mark.parametrize("n,d,K,order", [(30000, 64, 12, "random"), (30000, 64, 12, "by_class"),
                                         (20000, 100, 6, "random"), (12000, 128, 4, "by_class"),
                                         (9000, 33, 5, "random"), (20000, 8, 40, "by_class")])
def test_cov_over_nonzero_posteriors(orc, prec, n, d, K, order):
    """Separated mixtures started from the generating parameters: most
    posteriors are exactly zero and the covariance pass sums over the rest.
    Covariances within the form's tolerance of the oracle and of the dense
    sum over every point.  by_class orders the points by their largest
    posterior, so most segments list no point of a given component."""
    print(f"{NAME[prec]} cov over nonzero posteriors n={n} d={d} K={K} order={order}", flush=True)
    X, w, mu, cov = fem.gengmm(n, d, K, -40.0, 40.0, 0.05, 900 + d + K)
    U, ld = orc.factor_all(cov)
    if order == "by_class":
        _, R0 = orc.estep(X, w, mu, U, ld)
        X = np.ascontiguousarray(X[np.argsort(R0.argmax(1), kind="stable")])
    _, R_ref = orc.estep(X, w, mu, U, ld)
    _, _, cov_ref = orc.mstep(X, R_ref)
    R, (_, _, cov_s) = one_step(X, w, mu, cov, prec, dense=False)
    _, (_, _, cov_d) = one_step(X, w, mu, cov, prec, dense=True)
    zero = float(np.mean(R == 0))
    print(NAME[prec], "zero posteriors", zero, "cov vs oracle", rel(cov_s, cov_ref), "vs dense", rel(cov_s, cov_d),
          flush=True)
    assert zero > 0.6
    assert rel(cov_s, cov_ref) <= TOL[prec]["cov"]
    assert rel(cov_s, cov_d) <= TOL[prec]["cov"]
    for k in range(K):
        assert np.array_equal(cov_s[k], cov_s[k].T)


def test_bf16x3_d128_k128_per_iteration_errors(orc):
    """The d = 128, K = 128 precision case at a size the oracle finishes in
    seconds: both forms step from the same start (the generating mixture with
    perturbed means and inflated covariances), with covariance regularisation
    g = 0.01 on both sides since the smallest Dirichlet components hold fewer
    than d points at this n; the relative errors of LL, means and covariances
    are reported for every iteration."""
    n, d, K, iters, g = 32768, 128, 128, 4, 0.01
    print("bf16x3 d=128 K=128 trajectory", flush=True)
    X, w, mu, cov = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, 31)
    mu = mu + 0.3 * np.random.default_rng(3).standard_normal(mu.shape)
    cov = cov * 1.2
    w = np.full(K, 1.0 / K)
    rows = []
    with fem.Context(n, d, K, fem.PREC_BF16X3) as ctx:
        ctx.upload(X)
        ctx.set_regularization(g)
        wb, mub, covb = w, mu, cov
        wr, mur, covr = w, mu, cov
        for it in range(iters):
            U, ld = orc.factor_all(covr)
            ll_ref, R_ref = orc.estep(X, wr, mur, U, ld)
            wr, mur, covr = orc.mstep(X, R_ref)
            covr = orc.regularize(covr, g)
            ctx.set_params(wb, mub, covb)
            ctx.factor()
            ll = ctx.estep()
            wb, mub, covb = ctx.mstep()
            rows.append((abs(ll - ll_ref) / abs(ll_ref), rel(mub, mur), rel(covb, covr)))
            print(f"iter {it}: LL {rows[-1][0]:.2e} mu {rows[-1][1]:.2e} cov {rows[-1][2]:.2e}", flush=True)
    for ll_e, mu_e, cov_e in rows:
        assert ll_e <= 1e-4 and mu_e <= 1e-4 and cov_e <= 1e-4


def test_time_to_loglik_schedules_reach_the_fp64_target():
    """The time-to-loglik harness: a BF16x3 stage with periodic FP64 checks
    and an FP32 stage that hands over to FP64 both end on an FP64 E-step
    within 1e-6 of the converged FP64 log-likelihood.  Covariance
    regularisation g = 1e-3 (the reference's -g) keeps every form's
    covariances positive definite along the way."""
    from fastem_amd import ttl
    print("time-to-loglik schedules", flush=True)
    # a case where EM from k-means++ converges slowly (merged clusters) to a
    # well-conditioned optimum (condition numbers below 3e3; CPU oracle check)
    n, d, K = 40000, 16, 6
    X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, 323)[0]
    with fem.Context(n, d, K, fem.PREC_FP64, 0) as ctx:
        ctx.upload(X)
        ctx.set_regularization(1e-3)
        plain = ttl.run_plain(ctx, seed=1, tol=1e-6, cap=400, conv=1e-12)
        ll_star = plain["ll_final"]
        target = ll_star - 1e-6 * abs(ll_star)
        for stages in ([(fem.PREC_BF16X3, ttl.check_every(5))], [(fem.PREC_FP32, 1e-9)]):
            r = ttl.run_schedule(ctx, seed=1, ll_star=ll_star, stages=stages, tol=1e-6, cap=400)
            assert r["t_reach"] is not None
            t, ll, prec = r["trace"][-1]
            assert prec == fem.PREC_FP64 and ll >= target
