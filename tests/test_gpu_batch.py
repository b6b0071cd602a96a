"""Batched multi-fit engine (GPU): every fit of a batch of uneven (n, d, K)
matches the same fit run alone through the single-fit context (bit-identical
trajectory semantics: same LL, same stop, parameters within 1e-12), the graph
replay equals eager launches bit for bit, and the FP64 oracle agrees to 1e-10."""
import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu


def make_fits(seed, count):
    rng = np.random.default_rng(seed)
    fits = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e5))))
        d = int(rng.choice([8, 16, 32, 64]))
        K = int(rng.integers(4, 17))
        X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 1000 + i)[0]
        w, mu, cov = fem.kmeanspp_init(X, K, seed=i, lloyd_iters=5)
        fits.append((X, K, (w, mu, cov)))
    return fits


@pytest.fixture(scope="module")
def fits():
    return make_fits(7, 6)


def run_batch(fits, use_graph, tune=False, maxsweeps=30, eps=1e-8, also_eager=False, g=0.0):
    b = fem.Batch([(X.shape[0], X.shape[1], K) for X, K, _ in fits])
    for i, (X, K, init) in enumerate(fits):
        b.upload(i, X)
        b.set_params(i, *init)
    report = b.tune() if tune else None
    b.run(maxsweeps, eps=eps, g=g, use_graph=use_graph)
    res = [b.result(i) for i in range(len(fits))]
    eager = None
    if also_eager:
        # same batch, same tuned launch shape, kernels launched eagerly
        for i, (X, K, init) in enumerate(fits):
            b.set_params(i, *init)
        b.run(maxsweeps, eps=eps, use_graph=False)
        eager = [b.result(i) for i in range(len(fits))]
    b.close()
    return res, report, eager


def test_batch_graph_equals_eager_and_single(fits):
    print("batch graph vs eager vs single", flush=True)
    rg, report, re_ = run_batch(fits, use_graph=True, tune=True, also_eager=True)
    print(report, flush=True)
    for a, b in zip(rg, re_):
        assert a["ll"] == b["ll"] and a["iter_counter"] == b["iter_counter"]
        for key in ("w", "mu", "cov"):
            assert np.array_equal(a[key], b[key])
    for i, (X, K, init) in enumerate(fits):
        with fem.Context(X.shape[0], X.shape[1], K) as ctx:
            ctx.upload(X)
            it, ll, (w, mu, cov) = fem.train(ctx, *init, maxsweeps=30, eps=1e-8)
        r = rg[i]
        assert r["iter_counter"] == it, (i, r["iter_counter"], it)
        assert abs(r["ll"] - ll) <= 1e-12 * abs(ll)
        assert np.max(np.abs(r["cov"] - cov)) <= 1e-12 * np.max(np.abs(cov))
        assert r["status"] in (fem.Batch.CONVERGED, fem.Batch.SWEEP_CAP)


def test_batch_matches_oracle(orc, fits):
    print("batch vs oracle", flush=True)
    res, _, _ = run_batch(fits[:3], use_graph=True, maxsweeps=12, eps=1e-300)
    for (X, K, init), r in zip(fits[:3], res):
        ref = orc.train(X, *init, maxsweeps=12, eps=1e-300, svd=False)
        assert r["iter_counter"] == ref.rc
        assert abs(r["ll"] - ref.ll) <= 1e-10 * abs(ref.ll)
        assert np.max(np.abs(r["mu"] - ref.mu)) <= 1e-10 * np.max(np.abs(ref.mu))


def test_batch_regularized_matches_single_and_oracle(orc, fits):
    """g = 1e-3: the batch regularises the initial covariances before the
    first factorisation and after every M-step (EM2/EMAlgorithm.cpp:145,158),
    as the single-fit path and the oracle do."""
    print("batch g=1e-3", flush=True)
    g = 1e-3
    res, _, _ = run_batch(fits, use_graph=True, maxsweeps=15, eps=1e-300, g=g)
    for i, ((X, K, init), r) in enumerate(zip(fits, res)):
        with fem.Context(X.shape[0], X.shape[1], K) as ctx:
            ctx.upload(X)
            it, ll, (w, mu, cov) = fem.train(ctx, *init, maxsweeps=15, eps=1e-300, g=g)
        assert r["iter_counter"] == it
        assert abs(r["ll"] - ll) <= 1e-12 * abs(ll)
        assert np.max(np.abs(r["cov"] - cov)) <= 1e-12 * np.max(np.abs(cov))
        if i < 2:
            ref = orc.train(X, *init, maxsweeps=15, eps=1e-300, g=g, svd=False)
            assert abs(r["ll"] - ref.ll) <= 1e-10 * abs(ref.ll)
            assert np.max(np.abs(r["cov"] - ref.cov)) <= 1e-10 * np.max(np.abs(ref.cov))


def test_batch_config4_ranges(orc):
    """Fits at the edges of the many-fit workload's ranges (N up to 1e6,
    K up to 64, every d in {8, 16, 32, 64}) in one batch: each equals the fit
    run alone (1e-12) and the smaller ones the oracle (1e-10)."""
    shapes = [(1000000, 64, 64), (10000, 8, 4), (300000, 32, 40), (50000, 16, 64), (20000, 64, 7)]
    fits = []
    for i, (n, d, K) in enumerate(shapes):
        print(f"config4 fit n={n} d={d} K={K}", flush=True)
        X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, 4000 + i)[0]
        fits.append((X, K, fem.kmeanspp_init(X, K, seed=i, lloyd_iters=29)))
    res, report, _ = run_batch(fits, use_graph=True, tune=True, maxsweeps=6, eps=1e-300, g=1e-3)
    print(report, flush=True)
    for i, ((X, K, init), r) in enumerate(zip(fits, res)):
        with fem.Context(X.shape[0], X.shape[1], K) as ctx:
            ctx.upload(X)
            it, ll, (w, mu, cov) = fem.train(ctx, *init, maxsweeps=6, eps=1e-300, g=1e-3)
        # regularised EM need not increase the LL; a decrease stops the fit
        # (e = 0 < eps, EM2/EMAlgorithm.cpp:167-172) in both paths alike
        assert r["iter_counter"] == it and 2 <= it <= 7, (i, r["iter_counter"], it)
        assert abs(r["ll"] - ll) <= 1e-12 * abs(ll)
        for key, ref in (("w", w), ("mu", mu), ("cov", cov)):
            assert np.max(np.abs(r[key] - ref)) <= 1e-12 * np.max(np.abs(ref))
        if X.shape[0] <= 50000:
            o = orc.train(X, *init, maxsweeps=6, eps=1e-300, g=1e-3, svd=False)
            assert abs(r["ll"] - o.ll) <= 1e-10 * abs(o.ll)
            assert np.max(np.abs(r["cov"] - o.cov)) <= 1e-10 * np.max(np.abs(o.cov))


def test_batch_failure_is_isolated(fits):
    print("batch failure isolation", flush=True)
    X, K, (w, mu, cov) = fits[0]
    bad = cov.copy()
    bad[0] = -np.eye(X.shape[1])
    b = fem.Batch([(X.shape[0], X.shape[1], K)] * 2)
    b.upload(0, X)
    b.upload(1, X)
    b.set_params(0, w, mu, bad)
    b.set_params(1, w, mu, cov)
    b.run(10, eps=1e-300)
    assert b.result(0)["status"] == fem.Batch.FAILED
    r1 = b.result(1)
    assert r1["status"] == fem.Batch.SWEEP_CAP and r1["iter_counter"] == 11
    b.close()


def test_bic_k_sweep_prefers_true_order():
    """A K sweep on data drawn from a K=5 mixture scores lowest BIC at K=5."""
    print("BIC K sweep", flush=True)
    n, d, Ktrue = 60000, 8, 5
    X = fem.gengmm(n, d, Ktrue, -12.0, 12.0, 0.2, 99)[0]
    Ks = [2, 3, 4, 5, 6, 7, 8]
    b = fem.Batch([(n, d, K) for K in Ks])
    for i, K in enumerate(Ks):
        b.upload(i, X)
        b.set_params(i, *fem.kmeanspp_init(X, K, seed=3, lloyd_iters=29))
    b.run(300, eps=1e-10)
    scores = [fem.bic(b.result(i)["ll"], n, d, K) for i, K in enumerate(Ks)]
    b.close()
    print("BIC", dict(zip(Ks, scores)), flush=True)
    assert Ks[int(np.argmin(scores))] == Ktrue


def test_consecutive_batches_keep_every_fit():
    """Batches created one after another, each filled right after creation
    (the pattern of the learned initializer's evaluation): every fit of every
    batch, all on the same data and start, ends exactly like one run of the
    single-fit context.  The pools are cleared on the batch's own stream at
    creation; a clear queued on another stream could land after the first
    fits' uploads and wipe them."""
    n, d, K, fits = 65536, 32, 16, 96
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 10000256)[0]
    w0, mu0, cov0 = fem.kmeanspp_init(X, K, seed=3, lloyd_iters=29)
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        it1, ll1, (_, mu1, cov1) = fem.train(ctx, w0, mu0, cov0, maxsweeps=20, eps=1e-5)
    for rep in range(4):
        print(f"consecutive batch {rep}", flush=True)
        b = fem.Batch([(n, d, K)] * fits)
        for i in range(fits):
            b.upload(i, X)
            b.set_params(i, w0, mu0, cov0)
        b.run(20, eps=1e-5)
        res = [b.result(i) for i in range(fits)]
        b.close()
        for i, r in enumerate(res):
            assert r["status"] != fem.Batch.FAILED, (rep, i)
            assert r["ll"] == res[0]["ll"] and r["iter_counter"] == res[0]["iter_counter"], (rep, i)
            np.testing.assert_array_equal(r["mu"], res[0]["mu"])
        assert res[0]["iter_counter"] == it1
        assert abs(res[0]["ll"] - ll1) <= 1e-12 * abs(ll1)
        assert np.max(np.abs(res[0]["cov"] - cov1)) <= 1e-12 * np.max(np.abs(cov1))
