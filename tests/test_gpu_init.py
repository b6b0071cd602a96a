"""Initializers on the device against the FP64 oracle (GPU), through the C ABI
and through the eminit / em2 drivers.

Seeds and assignment counts are bit-exact (inputs without near-ties);
centroids and covariances agree to FP64 summation-order rounding (1e-12 /
1e-10 relative)."""
import re
import subprocess

import numpy as np
import pytest

import fastem_amd as fem
from fastem_amd import formats

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_mixture(n, got, ref):
    """Counts exact, means 1e-12, covariances 1e-10 (components with fewer
    than two points have no covariance: n_k - 1 <= 0 divides it)."""
    w, mu, cov = got
    w_ref, mu_ref, cov_ref = ref
    np.testing.assert_array_equal(w * n, w_ref * n)  # assignment counts, exact
    np.testing.assert_allclose(mu, mu_ref, rtol=1e-12, atol=1e-12)
    ok = np.rint(w_ref * n) >= 2
    assert np.max(np.abs(cov[ok] - cov_ref[ok])) <= 1e-10 * np.max(np.abs(cov_ref[ok]))


def device_init(X, K, seed, seeding, sweeps=29):
    n, d = X.shape
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        try:
            return ctx.kmeans_init(seed, seeding=seeding, max_sweeps=sweeps)
        except fem.FemError:
            return ctx.get_params()


@pytest.mark.parametrize("n,d,K,seed", [(30000, 6, 5, 9), (200000, 64, 64, 3), (50001, 17, 12, 4),
                                        (20000, 128, 32, 5)])
def test_forgy_lloyd_mixture(orc, n, d, K, seed):
    """KMeansEMInitializer (EM2/EMInitializer.cpp:82-134): Forgy rows, Lloyd,
    cluster mixture, all on one Rand() stream."""
    say(f"forgy n={n} d={d} K={K}")
    X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, seed)[0]
    got = device_init(X, K, seed, fem.SEED_FORGY)
    check_mixture(n, got, orc.kmeans_init(X, K, seed=seed))


def test_forgy_seeds_only(orc):
    """With no Lloyd sweep the means are the Forgy rows themselves."""
    say("forgy seeds")
    X = fem.gengmm(5000, 3, 4, -10.0, 10.0, 0.1, 5)[0]
    _, mu, _ = device_init(X, 4, 8, fem.SEED_FORGY, sweeps=0)
    np.testing.assert_array_equal(mu, orc.forgy(X, 4, orc.Rand(8)))


@pytest.mark.parametrize("n,d,K,seed", [(30000, 6, 5, 2), (100000, 32, 16, 7), (20000, 128, 8, 1),
                                        (200000, 64, 64, 11)])
def test_kmeansoror_seeds(orc, n, d, K, seed):
    """k-means|| candidates and the weighted k-means++ over them
    (Clust/KMeansOrOrInitializer.cpp:143-174): the seeds equal the oracle's
    one-thread run exactly."""
    say(f"kmeans|| seeds n={n} d={d} K={K}")
    X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, seed)[0]
    _, mu, _ = device_init(X, K, seed, fem.SEED_KMEANSOROR, sweeps=0)
    C, m = orc.kmeansoror(X, K, orc.Rand(seed))
    say("candidates", m)
    np.testing.assert_array_equal(mu, C)


@pytest.mark.parametrize("n,d,K,seed", [(30000, 6, 5, 2), (100000, 32, 16, 7)])
def test_kmeansoror_lloyd_mixture(orc, n, d, K, seed):
    say(f"kmeans|| mixture n={n} d={d} K={K}")
    X = fem.gengmm(n, d, K, -20.0, 20.0, 0.05, seed)[0]
    got = device_init(X, K, seed, fem.SEED_KMEANSOROR)
    check_mixture(n, got, orc.kmeansoror_init(X, K, seed=seed))


def test_empty_cluster_repair(orc):
    """A starting centroid far from every row owns no point after the first
    assignment, so Lloyd re-seeds it from a row drawn with Rand()
    (Clust/CentroidRepair.cpp:21-28); the repaired run equals the oracle's."""
    say("empty-cluster repair")
    n, d, K = 20000, 5, 6
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 17)[0]
    C0 = X[np.random.default_rng(0).choice(n, K, replace=False)].astype(np.float64)
    C0[2] = 1e6
    C0[4] = -1e6
    with fem.Context(n, d, K) as ctx:
        ctx.upload(X)
        got = ctx.kmeans_init_seeded(C0, seed=123, max_sweeps=29)
    rng = orc.Rand(123)
    C, _, _ = orc.lloyd(X, C0, rng)
    assert np.all(np.abs(C) < 1e3)  # both far centroids were repaired
    check_mixture(n, got, orc.gmm_from_clusters(X, C, orc.classify(X, C)))


def run_tool(path, args, timeout=300):
    p = subprocess.run([path] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.parametrize("name,banner", [("kmeans", "K-means EM initializer"),
                                         ("kmeanspp", "K-means++ EM initializer"),
                                         ("kmeansoror", "K-means|| EM initializer")])
def test_eminit_kmeans_family(orc, tmp_path, name, banner):
    """eminit -I <name> -r S -c K -o x.cmp (EM2/eminit.cpp): the written
    mixture equals the oracle's initializer on the same Rand() stream."""
    say(f"eminit -I {name}")
    n, d, K, seed = 40000, 12, 9, 6
    X = fem.gengmm(n, d, K, -15.0, 15.0, 0.1, 41)[0]
    formats.write_bin(tmp_path / "x.bin", X)
    rc, out = run_tool(fem.EMINIT_PATH, ["-I", name, "-r", seed, "-c", K, "-o", tmp_path / "x.cmp",
                                         tmp_path / "x.bin"])
    assert rc == 0, out
    assert banner in out
    got = formats.read_cmp(tmp_path / "x.cmp")
    ref = {"kmeans": orc.kmeans_init, "kmeanspp": orc.kmeanspp_init,
           "kmeansoror": orc.kmeansoror_init}[name](X, K, seed=seed)
    check_mixture(n, got, ref)


def test_eminit_default_is_kmeans(orc, tmp_path):
    say("eminit default")
    n, d, K = 8000, 4, 3
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 42)[0]
    formats.write_bin(tmp_path / "x.bin", X)
    rc, out = run_tool(fem.EMINIT_PATH, ["-r", 2, "-c", K, "-o", tmp_path / "x.cmp",
                                         tmp_path / "x.bin"])
    assert rc == 0 and "K-means EM initializer" in out, out
    check_mixture(n, formats.read_cmp(tmp_path / "x.cmp"), orc.kmeans_init(X, K, seed=2))


def test_em2_jain(orc, tmp_path):
    """em2 -I jain (the default initializer, EM2/EMInitializer.cpp:43-80):
    -m 0 writes the initial mixture itself; -m 6 follows the oracle's
    trajectory from orc_jain's mixture."""
    say("em2 -I jain")
    n, d, K, seed = 30000, 8, 5, 13
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 43)[0]
    formats.write_bin(tmp_path / "x.bin", X)
    w0, mu0, cov0 = orc.jain_init(X, K, orc.Rand(seed))
    rc, out = run_tool(fem.EM2_PATH, ["-c", K, "-r", seed, "-m", 0, "-S", "-o", tmp_path / "j.cmp",
                                      tmp_path / "x.bin"])
    assert rc == 0 and "Jain EM Initializer" in out, out
    w, mu, cov = formats.read_cmp(tmp_path / "j.cmp")
    np.testing.assert_array_equal(w, w0)
    np.testing.assert_array_equal(mu, mu0)
    np.testing.assert_allclose(cov, cov0, rtol=1e-12, atol=0)
    rc, out = run_tool(fem.EM2_PATH, ["-c", K, "-r", seed, "-m", 6, "-f", "1e-300", "-S", "-v", 3,
                                      tmp_path / "x.bin"])
    assert rc == 0, out
    trace = np.array([float(v) for v in re.findall(r"LogL\(17g\): (\S+)", out)])
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=6, eps=1e-300, svd=False)
    np.testing.assert_allclose(trace, ref.ll_trace, rtol=1e-10)
