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


@pytest.mark.parametrize("header", ["fastem_abi.h", "fastem_tools.h", "fastem_setnet.h"])
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

