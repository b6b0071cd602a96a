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
