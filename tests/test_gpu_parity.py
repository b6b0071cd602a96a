"""Parity of the HIP engine (through the C ABI) with the FP64 oracle.

Plain FP64 form: per-iteration log-likelihood, weights, means and covariances
within 1e-10 relative of the oracle on the same data and the same initial
parameters.  Needs an MI355X.
"""
import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu

LL_RTOL = 1e-10
PARAM_RTOL = 1e-10


def say(*a):
    print(*a, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def mixture(n, d, K, seed, lo=-10.0, hi=10.0, diag=0.1):
    X, w, mu, cov = fem.gengmm(n, d, K, lo, hi, diag, seed)
    return X, w, mu, cov


def jittered_init(X, K, seed):
    """Initial parameters: K distinct data rows as means, the data covariance
    for every component, equal weights."""
    rng = np.random.default_rng(seed)
    n, d = X.shape
    idx = rng.choice(n, size=K, replace=False)
    mu = X[idx].astype(np.float64)
    C = np.cov(X.astype(np.float64), rowvar=False).reshape(d, d) + 1e-3 * np.eye(d)
    return np.full(K, 1.0 / K), mu, np.stack([C] * K)


@pytest.mark.parametrize("d", [1, 3, 16, 17, 64, 128])
def test_factor_logdet(orc, d):
    say(f"factor d={d}")
    K = 7
    rng = np.random.default_rng(d)
    A = rng.standard_normal((K, d, d))
    cov = A @ A.transpose(0, 2, 1) / d + 0.1 * np.eye(d)
    with fem.Context(100, d, K) as ctx:
        ctx.set_params(np.full(K, 1.0 / K), rng.standard_normal((K, d)), cov)
        ld, info = ctx.factor()
    assert np.all(info == 0)
    _, ld_ref = orc.factor_all(cov)
    assert rel_err(ld, ld_ref) < 1e-12


def test_factor_not_pd(orc):
    say("factor not-PD")
    d, K = 4, 3
    cov = np.stack([np.eye(d)] * K)
    cov[1, 2, 2] = -1.0
    with fem.Context(64, d, K) as ctx:
        ctx.set_params(np.full(K, 1.0 / K), np.zeros((K, d)), cov)
        with pytest.raises(fem.FemError) as ei:
            ctx.factor()
        assert ei.value.code == fem.FEM_ERR_NOT_PD
        assert list(ei.value.info) == [0, 3, 0]
        with pytest.raises(fem.FemError) as ei:
            ctx.estep()  # no E-step on a failed factorisation
        assert ei.value.code == fem.FEM_ERR_STATE


def test_staged_params_download():
    """fem_get_params_start / _finish return the parameters as they were when
    the copies were queued, with later device work already behind them."""
    say("staged parameter download")
    n, d, K = 20000, 10, 5
    X, _, _, _ = mixture(n, d, K, seed=77)
    w, mu, cov = jittered_init(X, K, seed=1)
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        ctx.set_params(w, mu, cov)
        with pytest.raises(fem.FemError):
            ctx.get_params_finish()
        ctx.factor()
        ctx.estep()
        p1 = ctx.mstep()
        ctx.get_params_start()
        ctx.factor()
        ctx.estep()
        p2 = ctx.mstep()
        s1 = ctx.get_params_finish()
        with pytest.raises(fem.FemError):
            ctx.get_params_finish()
        ctx.get_params_start()
        s2 = ctx.get_params_finish()
    for a, b, c, e in zip(p1, s1, p2, s2):
        np.testing.assert_array_equal(b, a)
        np.testing.assert_array_equal(e, c)
    assert not np.array_equal(p1[1], p2[1])


@pytest.mark.parametrize("n,d,K", [(1000, 3, 2), (5003, 17, 7), (4096, 16, 8), (3000, 64, 5),
                                   (2500, 128, 3), (777, 1, 1), (9000, 8, 64), (4099, 2, 16),
                                   (3001, 90, 4), (2600, 100, 3), (2200, 112, 2), (6000, 12, 100),
                                   (2800, 72, 3), (2300, 80, 6), (2400, 97, 3), (2100, 120, 2)])
def test_estep_mstep_parity(orc, n, d, K):
    say(f"estep/mstep n={n} d={d} K={K}")
    X, _, _, _ = mixture(n, d, K, seed=n + d + K)
    w, mu, cov = jittered_init(X, K, seed=1)
    U, ld = orc.factor_all(cov)
    ll_ref, R_ref = orc.estep(X, w, mu, U, ld)
    w_ref, mu_ref, cov_ref = orc.mstep(X, R_ref)
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        ctx.set_params(w, mu, cov)
        ctx.factor()
        ll = ctx.estep()
        R = ctx.posteriors()
        w2, mu2, cov2 = ctx.mstep()
    assert abs(ll - ll_ref) <= LL_RTOL * abs(ll_ref)
    np.testing.assert_allclose(R, R_ref, rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(R.sum(1), 1.0, rtol=1e-13)
    assert rel_err(w2, w_ref) < PARAM_RTOL
    assert rel_err(mu2, mu_ref) < PARAM_RTOL
    assert rel_err(cov2, cov_ref) < PARAM_RTOL
    for k in range(K):
        assert np.array_equal(cov2[k], cov2[k].T)


def mstep_with(X, w, mu, cov, dense):
    """Posteriors and M-step of one FP64 iteration; dense=True sums the
    covariances over every point (FASTEM_COV_DENSE=1), else over the points
    with a nonzero posterior only (the default)."""
    import os
    n, d = X.shape
    if dense:
        os.environ["FASTEM_COV_DENSE"] = "1"
    try:
        with fem.Context(n, d, len(w)) as ctx:
            ctx.upload(X)
            ctx.set_params(w, mu, cov)
            ctx.factor()
            ctx.estep()
            R = ctx.posteriors()
            return R, ctx.mstep()
    finally:
        os.environ.pop("FASTEM_COV_DENSE", None)


@pytest.mark.parametrize("n,d,K,order", [(30000, 16, 8, "random"), (30000, 64, 12, "random"),
                                         (20000, 100, 6, "random"), (12000, 128, 4, "random"),
                                         (30000, 64, 12, "by_class"), (9000, 33, 5, "by_class")])
def test_cov_over_nonzero_posteriors(orc, n, d, K, order):
    """Separated mixtures (means U[-40, 40], started from the generating
    parameters): most posteriors are exactly zero and the covariance pass
    sums over the rest.  Covariances within 1e-10 of the oracle and 1e-13 of
    the dense sum over every point.  by_class orders the points by their
    largest posterior, so most 1024-point segments list no point of a given
    component and the rest list every point."""
    say(f"cov over nonzero posteriors n={n} d={d} K={K} order={order}")
    X, w, mu, cov = fem.gengmm(n, d, K, -40.0, 40.0, 0.05, 700 + d + K)
    U, ld = orc.factor_all(cov)
    if order == "by_class":
        _, R0 = orc.estep(X, w, mu, U, ld)
        X = np.ascontiguousarray(X[np.argsort(R0.argmax(1), kind="stable")])
    ll_ref, R_ref = orc.estep(X, w, mu, U, ld)
    w_ref, mu_ref, cov_ref = orc.mstep(X, R_ref)
    R, (w2, mu2, cov2) = mstep_with(X, w, mu, cov, dense=False)
    _, (_, _, cov_dense) = mstep_with(X, w, mu, cov, dense=True)
    zero = np.mean(R == 0)
    say("zero posteriors", zero, "cov rel err vs oracle", rel_err(cov2, cov_ref),
        "vs dense", rel_err(cov2, cov_dense))
    assert zero > 0.6
    assert rel_err(w2, w_ref) < PARAM_RTOL
    assert rel_err(mu2, mu_ref) < PARAM_RTOL
    assert rel_err(cov2, cov_ref) < PARAM_RTOL
    assert rel_err(cov2, cov_dense) < 1e-13
    for k in range(K):
        assert np.array_equal(cov2[k], cov2[k].T)


@pytest.mark.parametrize("g", [0.0, 0.01])
def test_trajectory_config1(orc, g):
    """N=100,000, d=16, K=8, means U[-10,10], cov AA^T/d + 0.1 I, k-means++
    init, 100 iterations; every iteration's LL and the final parameters
    within 1e-10."""
    n, d, K, iters = 100000, 16, 8, 100
    say(f"trajectory config1 g={g}")
    X, _, _, _ = mixture(n, d, K, seed=11)
    w0, mu0, cov0 = fem.kmeanspp_init(X, K, seed=5)
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, g=g, svd=False)
    assert ref.rc == iters + 1
    trace = []
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        it, ll, (w, mu, cov) = fem.train(ctx, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, g=g,
                                         ll_trace=trace)
    assert it == ref.rc
    errs = np.abs(np.array(trace) - ref.ll_trace) / np.abs(ref.ll_trace)
    say("max per-iteration LL rel err", errs.max())
    assert errs.max() < LL_RTOL
    assert rel_err(w, ref.w) < PARAM_RTOL
    assert rel_err(mu, ref.mu) < PARAM_RTOL
    assert rel_err(cov, ref.cov) < PARAM_RTOL


def run_dense_or_pruned(X, w0, mu0, cov0, iters, dense):
    """fem.train for `iters` iterations with the FP64 E-step evaluating every
    component in point order (dense=True, FASTEM_ESTEP_DENSE=1) or pruned
    over the sorted slot order (the default); returns the LL trace, the
    parameters, the last E-step's posteriors and classes (point order) and
    the share of (wave, component) evaluations the last E-step skipped."""
    import os
    n, d = X.shape
    K = len(w0)
    if dense:
        os.environ["FASTEM_ESTEP_DENSE"] = "1"
    try:
        trace = []
        with fem.Context(n, d, K) as ctx:
            ctx.upload(X)
            _, _, params = fem.train(ctx, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, ll_trace=trace)
            return trace, params, ctx.posteriors(), ctx.classes(), ctx.estep_skipped()
    finally:
        os.environ.pop("FASTEM_ESTEP_DENSE", None)


@pytest.mark.parametrize("n,d,K,lo,hi,iters", [(60000, 40, 12, -20.0, 20.0, 25), (60000, 64, 16, -20.0, 20.0, 8),
                                               (30000, 17, 5, -10.0, 10.0, 20), (20000, 100, 6, -20.0, 20.0, 10),
                                               (80000, 32, 40, -20.0, 20.0, 6)])
def test_pruned_estep_trajectory(orc, n, d, K, lo, hi, iters):
    """The pruned FP64 E-step (points re-sorted by their component after
    E-steps 1, 2, 4, ...; components skipped per wave on the 16-row bound)
    against the oracle's TrainNoCopy from the generating mixture, perturbed: per-iteration LL
    and final parameters within 1e-10; and the LL trace, last posteriors and
    classes (which leave the context in point order) against the same run
    with every component evaluated in point order."""
    say(f"pruned trajectory n={n} d={d} K={K} iters={iters}")
    X, w, mu, cov = fem.gengmm(n, d, K, lo, hi, 0.05, 500 + d + K)
    rng = np.random.default_rng(d + K)
    w0, mu0, cov0 = w, mu + 0.3 * rng.standard_normal(mu.shape), 1.2 * cov
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, svd=False)
    trace, (w, mu, cov), R, cls, skipped = run_dense_or_pruned(X, w0, mu0, cov0, iters, dense=False)
    trace_d, (w_d, mu_d, cov_d), R_d, cls_d, skipped_d = run_dense_or_pruned(X, w0, mu0, cov0, iters, dense=True)
    errs = np.abs(np.array(trace) - ref.ll_trace) / np.abs(ref.ll_trace)
    say("skipped", skipped, "max LL rel err vs oracle", errs.max(),
        "vs dense", np.max(np.abs(np.array(trace) - trace_d) / np.abs(trace_d)))
    assert skipped_d == 0.0
    assert errs.max() < LL_RTOL
    assert rel_err(w, ref.w) < PARAM_RTOL
    assert rel_err(mu, ref.mu) < PARAM_RTOL
    assert rel_err(cov, ref.cov) < PARAM_RTOL
    assert np.max(np.abs(np.array(trace) - trace_d) / np.abs(trace_d)) < 1e-12
    np.testing.assert_allclose(R, R_d, rtol=1e-9, atol=1e-13)
    assert np.mean(cls == cls_d) > 0.9999
    np.testing.assert_allclose(R.sum(1), 1.0, rtol=1e-13)
    if d >= 32:
        assert skipped > 0.5


def test_pruned_classes_follow_the_points(orc):
    """Points in a random order and in class order give the same per-point
    posteriors and classes through the sorted slot order, and the classes
    are the argmax of the oracle's posteriors."""
    n, d, K = 30000, 48, 10
    say("pruned classes in point order")
    X, w, mu, cov = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, 77)
    perm = np.random.default_rng(5).permutation(n)
    outs = []
    for Xo in (X, np.ascontiguousarray(X[perm])):
        with fem.Context(n, d, K) as ctx:
            ctx.upload(Xo)
            ctx.set_params(w, mu, cov)
            for _ in range(3):  # sorted after the first E-step
                ctx.factor()
                ctx.estep()
            outs.append((ctx.posteriors(), ctx.classes(), ctx.estep_skipped()))
    (R0, c0, s0), (R1, c1, s1) = outs
    say("skipped", s0, s1)
    inv = np.argsort(perm)
    np.testing.assert_allclose(R1[inv], R0, rtol=1e-12, atol=1e-15)
    assert np.array_equal(c1[inv], c0)
    U, ld = orc.factor_all(cov)
    _, R_ref = orc.estep(X, w, mu, U, ld)
    np.testing.assert_allclose(R0, R_ref, rtol=1e-9, atol=1e-13)
    assert np.array_equal(c0, R_ref.argmax(1))
    assert s0 > 0.5 and s1 > 0.5


def test_convergence_stop_matches(orc):
    n, d, K = 20000, 5, 4
    say("convergence stop")
    X, _, _, _ = mixture(n, d, K, seed=3)
    w0, mu0, cov0 = jittered_init(X, K, seed=2)
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=500, eps=1e-7, svd=False)
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        it, ll, _ = fem.train(ctx, w0, mu0, cov0, maxsweeps=500, eps=1e-7)
    assert it == ref.rc
    assert abs(ll - ref.ll) <= LL_RTOL * abs(ref.ll)


def test_posteriors_deterministic():
    n, d, K = 50000, 16, 8
    say("determinism")
    X, w, mu, cov = mixture(n, d, K, seed=4)
    outs = []
    for _ in range(2):
        with fem.Context(n, d, K) as ctx:
            ctx.upload(X)
            ctx.set_params(w, mu, cov)
            ctx.factor()
            ll = ctx.estep()
            p = ctx.mstep()
            outs.append((ll, p))
    assert outs[0][0] == outs[1][0]
    for a, b in zip(outs[0][1], outs[1][1]):
        assert np.array_equal(a, b)


def test_full_size_properties(orc):
    """N=4,194,304, d=64, K=64 (means U[-20,20], cov AA^T/d + 0.05 I): the LL of
    the whole set equals the sum over two halves run separately, posteriors
    of sampled rows equal the oracle's for those rows, weights sum to one and
    the LL does not decrease."""
    n, d, K = 4194304, 64, 64
    say("full size generate")
    X, w, mu, cov = mixture(n, d, K, seed=2024, lo=-20.0, hi=20.0, diag=0.05)
    rng = np.random.default_rng(0)
    mu0 = mu + 0.5 * rng.standard_normal(mu.shape)
    cov0 = cov * 1.3
    w0 = np.full(K, 1.0 / K)
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        ctx.set_params(w0, mu0, cov0)
        ctx.factor()
        say("full size estep")
        ll_full = ctx.estep()
        R = ctx.posteriors()
        sample = rng.choice(n, 512, replace=False)
        R_s = R[sample]
        del R
        lls = [ll_full]
        for _ in range(3):
            w1, mu1, cov1 = ctx.mstep()
            assert abs(w1.sum() - 1.0) < 1e-12
            ctx.factor()
            lls.append(ctx.estep())
    assert np.all(np.diff(lls) >= -1e-12 * np.abs(lls[1:]))
    halves = []
    for a, b in ((0, n // 2), (n // 2, n)):
        say("full size half", a)
        with fem.Context(b - a, d, K) as ctx:
            ctx.upload(X[a:b])
            ctx.set_params(w0, mu0, cov0)
            ctx.factor()
            halves.append(ctx.estep())
    assert abs(sum(halves) - ll_full) <= 1e-12 * abs(ll_full)
    U, ld = orc.factor_all(cov0)
    _, R_ref = orc.estep(X[sample], w0, mu0, U, ld)
    np.testing.assert_allclose(R_s, R_ref, rtol=1e-9, atol=1e-13)


@pytest.mark.timeout(900)
def test_config2_full_size_vs_oracle(orc):
    """The benchmark workload at full size against the oracle: N=4,194,304,
    d=64, K=64 (means U[-20,20], cov AA^T/d + 0.05 I), device k-means++ +
    Lloyd init, then 3 EM iterations of the FP64 form and of the oracle's
    TrainNoCopy restatement from that identical init.  Every iteration's LL
    and the final weights, means and covariances agree within 1e-10."""
    n, d, K, iters = 4194304, 64, 64, 3
    say("config2 generate")
    X, _, _, _ = mixture(n, d, K, seed=2025, lo=-20.0, hi=20.0, diag=0.05)
    say("config2 device kmeans++ init")
    w0, mu0, cov0 = fem.kmeanspp_init(X, K, seed=1)
    trace = []
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        say("config2 device EM")
        it, ll, (w, mu, cov) = fem.train(ctx, w0, mu0, cov0, maxsweeps=iters, eps=1e-300,
                                         ll_trace=trace)
    say("config2 oracle EM (host, FP64)")
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, svd=False)
    assert ref.rc == iters + 1 and it == ref.rc
    ll_err = np.abs(np.array(trace) - ref.ll_trace) / np.abs(ref.ll_trace)
    errs = {"ll": ll_err.max(), "w": rel_err(w, ref.w), "mu": rel_err(mu, ref.mu),
            "cov": rel_err(cov, ref.cov)}
    say("config2 max rel errors vs oracle:", errs, "LL trace", list(ref.ll_trace))
    assert errs["ll"] < LL_RTOL
    assert errs["w"] < PARAM_RTOL and errs["mu"] < PARAM_RTOL and errs["cov"] < PARAM_RTOL


@pytest.mark.parametrize("prec", [fem.PREC_FP64, fem.PREC_FP32])
def test_unstored_log_densities_have_zero_posteriors(orc, prec):
    """The E-steps store a log-density only while it is within the
    exponential's underflow range of the point's running maximum (760 in
    FP64, 110 in FP32).  A 1-d mixture puts components about 741 and 780
    (FP64) or 100 and 120 (FP32) below a point's own: posteriors against the
    FP64 oracle, exactly zero where the oracle's are, whatever order the
    components come in (weights in both orders)."""
    print(f"unstored log-densities precision={prec}", flush=True)
    far = (38.5, 39.5) if prec == fem.PREC_FP64 else (14.1, 15.5)
    rng = np.random.default_rng(3)
    X = np.concatenate([rng.normal(0.0, 0.3, 3000), rng.normal(far[0], 0.3, 500),
                        rng.normal(far[1], 0.3, 500), rng.normal(-far[0], 0.3, 96)])
    X = X.astype(np.float32)[:, None]
    mu = np.array([[0.0], [far[0]], [far[1]], [-far[0]]])
    cov = np.ones((4, 1, 1))
    for w in (np.array([0.7, 0.1, 0.1, 0.1]), np.array([0.05, 0.15, 0.3, 0.5])):
        with fem.Context(X.shape[0], 1, 4, prec) as ctx:
            ctx.upload(X)
            ctx.set_params(w, mu, cov)
            ctx.factor()
            ll = ctx.estep()
            R = ctx.posteriors()
        U, ld = orc.factor_all(cov)
        ll_ref, R_ref = orc.estep(X, w, mu, U, ld)
        # log-density gaps to the point's log-sum-exp (1-d, unit variances)
        s = np.log(w)[None, :] - 0.5 * np.log(2 * np.pi) - 0.5 * (X.astype(np.float64) - mu[:, 0][None, :]) ** 2
        gap = s - np.logaddexp.reduce(s, axis=1)[:, None]
        rng_alive = 760.0 if prec == fem.PREC_FP64 else 110.0
        unstored = gap < -rng_alive - 1.0  # below the range even against the final maximum
        assert unstored.any() and ((gap > -rng_alive + 1.0) & (gap < -rng_alive + 30.0)).any()
        np.testing.assert_array_equal(R[unstored], 0.0)
        if prec == fem.PREC_FP64:
            np.testing.assert_array_equal(R_ref[unstored], 0.0)  # FP64 exp underflows there too
        if prec == fem.PREC_FP64:
            assert abs(ll - ll_ref) <= 1e-12 * abs(ll_ref)
            np.testing.assert_allclose(R, R_ref, rtol=1e-9, atol=1e-320)
        else:
            assert abs(ll - ll_ref) <= 1e-6 * abs(ll_ref)
            np.testing.assert_allclose(R, R_ref, rtol=1e-3, atol=1e-30)
