"""Device k-means initialisation against the FP64 oracle (GPU).

Seeds and assignment counts must be bit-exact (inputs without near-ties);
centroids and covariances agree to FP64 summation-order rounding."""
import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu


def test_kmeanspp_seeding_only(orc):
    print("kmeans++ seeding", flush=True)
    X = fem.gengmm(5000, 3, 4, -10.0, 10.0, 0.1, 5)[0]
    w, mu, cov = fem.kmeanspp_init(X, 4, seed=8, lloyd_iters=0)
    C = orc.kmeanspp(X, 4, orc.Rand(8))
    np.testing.assert_array_equal(mu, C)
    cls = orc.classify(X, C)
    np.testing.assert_array_equal(w, np.bincount(cls, minlength=4) / 5000)


@pytest.mark.parametrize("n,d,K,seed", [(30000, 6, 5, 9), (200000, 64, 64, 3), (50001, 17, 12, 4),
                                        (20000, 128, 32, 5)])
def test_kmeanspp_lloyd_mixture(orc, n, d, K, seed):
    print(f"kmeans++ + Lloyd n={n} d={d} K={K}", flush=True)
    X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, seed)[0]
    w, mu, cov = fem.kmeanspp_init(X, K, seed=seed, lloyd_iters=29)
    w_ref, mu_ref, cov_ref = orc.kmeanspp_init(X, K, seed=seed)
    np.testing.assert_array_equal(w * n, w_ref * n)  # assignment counts, exact
    np.testing.assert_allclose(mu, mu_ref, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(cov - cov_ref)) <= 1e-10 * np.max(np.abs(cov_ref))


def test_kmeans_init_on_context_sets_params(orc):
    print("kmeans init on context", flush=True)
    n, d, K = 40000, 8, 6
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 12)[0]
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        w, mu, cov = ctx.kmeans_init(seed=2, seeding=1, max_sweeps=29)
        w2, mu2, cov2 = ctx.get_params()
        assert np.array_equal(w, w2) and np.array_equal(mu, mu2) and np.array_equal(cov, cov2)
        ctx.factor()
        ll = ctx.estep()
        wf, muf, covf = ctx.kmeans_init(seed=2, seeding=0, max_sweeps=29)  # Forgy
        assert abs(wf.sum() - 1.0) < 1e-12
    U, ld = orc.factor_all(cov)
    ll_ref, _ = orc.estep(X, w, mu, U, ld)
    assert abs(ll - ll_ref) <= 1e-10 * abs(ll_ref)


def grid_points(n, d, seed):
    """n rows drawn from the integer grid {-2..2}^d (many rows equidistant
    from symmetric centroids)."""
    rng = np.random.default_rng(seed)
    return rng.integers(-2, 3, size=(n, d)).astype(np.float32)


@pytest.mark.parametrize("case", ["exact_ties", "near_ties", "random_d64_k64", "separated", "single"])
def test_assignment_matches_oracle_classification(orc, case):
    """The assignment pass (distances' cross terms on FP64 MFMA, the exact
    sequential distance wherever the MFMA gap does not prove the argmin)
    against the oracle's classification by the same centroids: exact ties
    (lowest index wins), near-ties 1e-15 apart (below the MFMA error bound,
    decided by the exact distances), a random d = 64, K = 64 case, and
    the generating means of well-separated clusters, and a single centroid.
    Zero Lloyd sweeps: the mixture is that of the
    classification by C0; counts must be equal and covariances within
    1e-12."""
    print(f"assignment vs oracle: {case}", flush=True)
    if case in ("exact_ties", "near_ties"):
        n, d = 40000, 4
        X = grid_points(n, d, 3)
        C0 = np.array([[1, 0, 0, 0], [-1, 0, 0, 0], [0, 1, 0, 0], [0, -1, 0, 0], [0, 0, 1, 1],
                       [0, 0, -1, -1]], dtype=np.float64)
        if case == "near_ties":
            C0[1, 0] -= 1e-15
            C0[3, 1] += 3e-16
            C0[5, 2] -= 2e-15
    elif case == "random_d64_k64":
        n, d, K = 100000, 64, 64
        X = fem.gengmm(n, d, K, -5.0, 5.0, 0.5, 41)[0]
        C0 = X[np.random.default_rng(2).choice(n, K, replace=False)].astype(np.float64)
    elif case == "separated":
        n, d, K = 30000, 17, 9
        X, _, C0, _ = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 42)
        C0 = np.ascontiguousarray(C0, dtype=np.float64)
    else:  # one centroid: no second distance, every argmin proven
        n, d = 3001, 3
        X = fem.gengmm(n, d, 2, -10.0, 10.0, 0.1, 43)[0]
        C0 = np.ascontiguousarray(X[:1], dtype=np.float64)
    K = C0.shape[0]
    cls = orc.classify(X, C0)
    assert np.bincount(cls, minlength=K).min() >= 2
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        w, mu, cov = ctx.kmeans_init_seeded(C0, seed=1, max_sweeps=0)
    w_ref, mu_ref, cov_ref = orc.gmm_from_clusters(X, C0, cls)
    np.testing.assert_array_equal(np.round(w * n), np.bincount(cls, minlength=K))
    np.testing.assert_array_equal(mu, C0)
    assert np.max(np.abs(cov - cov_ref)) <= 1e-12 * np.max(np.abs(cov_ref))


@pytest.mark.parametrize("n,d,K,seed", [(30000, 6, 5, 9), (200000, 64, 64, 3), (50001, 17, 12, 4)])
def test_kmeanspp_device_draw_equals_host_draw(monkeypatch, n, d, K, seed):
    """The k-means++ D^2 draw decided on the device (prefix sums in any
    order, a rounding bound that proves the sequential rule's row) against
    the host's sequential FP64 prefix sums (FASTEM_KPP_DRAW=host, the path
    the device falls back to): the same seeds, hence the same mixture."""
    print(f"kmeans++ device vs host draw n={n} d={d} K={K}", flush=True)
    X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, seed)[0]
    dev = fem.kmeanspp_init(X, K, seed=seed, lloyd_iters=0)
    monkeypatch.setenv("FASTEM_KPP_DRAW", "host")
    host = fem.kmeanspp_init(X, K, seed=seed, lloyd_iters=0)
    for a, b in zip(dev, host):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n,d,K,seed", [(20000, 8, 12, 21), (30001, 17, 6, 22), (12000, 64, 40, 23)])
def test_greedy_kmeanspp_seeds_match_replay(orc, n, d, K, seed):
    """Greedy k-means++ (the batched workload's seeding): the seeds equal a
    numpy replay on the same Rand() stream (candidate draws by the k-means++
    rule, least potential kept); then Lloyd and the mixture as k-means++."""
    print(f"greedy kmeans++ n={n} d={d} K={K}", flush=True)
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, seed)[0]
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        w, mu, cov = ctx.kmeans_init(seed, seeding=fem.SEED_KMEANSPP_GREEDY, max_sweeps=0)
    C = orc.kmeanspp_greedy(X, K, orc.Rand(seed))
    np.testing.assert_array_equal(mu, C)  # no Lloyd sweep: the means are the seeds
    cls = orc.classify(X, C)
    np.testing.assert_array_equal(np.rint(w * n), np.bincount(cls, minlength=K))  # w = n_k / n
    _, _, cov_ref = orc.gmm_from_clusters(X, C, cls)
    assert np.max(np.abs(cov - cov_ref)) <= 1e-10 * np.max(np.abs(cov_ref))


def test_greedy_kmeanspp_lowers_the_potential():
    """On well-separated mixtures the greedy seeding leaves a lower k-means
    potential (sum of squared distances to the nearest seed) than k-means++
    from the same seeds of the stream, over ten streams."""
    print("greedy kmeans++ potential", flush=True)
    n, d, K = 40000, 16, 32
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 31)[0]
    X64 = X.astype(np.float64)

    def potential(C):
        dd = ((X64[:, None, :] - C[None, :, :]) ** 2).sum(-1)
        return dd.min(1).sum()

    wins = 0
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        for s in range(10):
            g = ctx.kmeans_init(100 + s, seeding=fem.SEED_KMEANSPP_GREEDY, max_sweeps=0)[1]
            p = ctx.kmeans_init(100 + s, seeding=fem.SEED_KMEANSPP, max_sweeps=0)[1]
            wins += potential(g) < potential(p)
    assert wins >= 8, wins
