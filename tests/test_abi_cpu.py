"""CPU-only checks of the product library: it loads without a GPU, exports every
symbol its public headers declare, reports errors without touching a device,
and its host tools (generator, k-means++ initializer) behave as specified."""
import hashlib
import os
import re

import numpy as np
import pytest

import fastem_amd as fem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fem_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["fastem_abi.h", "fastem_tools.h"])
def test_library_exports_every_declared_symbol(header):
    L = fem.lib()
    names = declared(header)
    assert names, header
    for name in names:
        assert hasattr(L, name), f"{name} declared in {header} but not exported"


def test_version_and_errors_without_gpu():
    assert b"gfx950" in fem.lib().fem_version()
    # argument errors are reported before any device is touched
    with pytest.raises(fem.FemError) as ei:
        fem.Context(0, 4, 2)
    assert ei.value.code == fem.FEM_ERR_ARG
    with pytest.raises(fem.FemError) as ei:
        fem.Context(100, 129, 2)
    assert "128" in str(ei.value)


def test_generator_deterministic_and_statistics():
    a = fem.gengmm(20000, 5, 3, -10.0, 10.0, 0.1, 42, labels=True)
    b = fem.gengmm(20000, 5, 3, -10.0, 10.0, 0.1, 42, labels=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    X, w, mu, cov, lab = a
    assert abs(w.sum() - 1.0) < 1e-15 and np.all(w > 0)
    assert np.all((mu >= -10) & (mu <= 10))
    for k in range(3):
        sel = X[lab == k].astype(np.float64)
        assert abs(len(sel) / 20000 - w[k]) < 0.02
        np.testing.assert_allclose(sel.mean(0), mu[k], atol=0.1)
        np.testing.assert_allclose(np.cov(sel, rowvar=False), cov[k], atol=0.15)
        assert np.all(np.linalg.eigvalsh(cov[k]) >= 0.1 - 1e-12)
    c = fem.gengmm(20000, 5, 3, -10.0, 10.0, 0.1, 43)[0]
    assert not np.array_equal(c, X)


def test_generator_fixed_bytes():
    # pins the generator's output stream for (n=1000, d=4, K=3, seed=7)
    X = fem.gengmm(1000, 4, 3, -10.0, 10.0, 0.1, 7)[0]
    h = hashlib.sha256(X.tobytes()).hexdigest()
    ref = open(os.path.join(ROOT, "tests", "golden", "gengmm_1000_4_3_7.sha256")).read().split()[0]
    assert h == ref


def test_kmeanspp_init_matches_oracle(orc):
    X = fem.gengmm(30000, 6, 5, -10.0, 10.0, 0.1, 9)[0]
    w, mu, cov = fem.kmeanspp_init(X, 5, seed=3, lloyd_iters=29)
    rng = orc.Rand(3)
    C = orc.kmeanspp(X, 5, rng)
    C, _, _ = orc.lloyd(X, C, rng, minrel=1e-4, maxiter=30)
    cls = orc.classify(X, C)
    w_ref, mu_ref, cov_ref = orc.gmm_from_clusters(X, C, cls)
    np.testing.assert_array_equal(w, w_ref)  # counts are exact
    np.testing.assert_allclose(mu, mu_ref, rtol=1e-12)
    np.testing.assert_allclose(cov, cov_ref, rtol=1e-10)


def test_kmeanspp_seeding_only(orc):
    X = fem.gengmm(5000, 3, 4, -10.0, 10.0, 0.1, 5)[0]
    w, mu, cov = fem.kmeanspp_init(X, 4, seed=8, lloyd_iters=0)
    C = orc.kmeanspp(X, 4, orc.Rand(8))
    np.testing.assert_array_equal(mu, C)
    cls = orc.classify(X, C)
    np.testing.assert_array_equal(w, np.bincount(cls, minlength=4) / 5000)
