"""Host-side parts of the reference's class API (CPU): the condition number
em2 checks every iteration unless -S is given, the Bhattacharyya distance,
and (GPU) the class of each row from the last E-step."""
import numpy as np
import pytest

import fastem_amd as fem


def spd_with_spectrum(ev, seed):
    """Q diag(ev) Q^T with a random orthogonal Q: the condition number is
    max(ev) / min(ev) exactly, up to the rounding of the product."""
    d = len(ev)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((d, d)))
    return (Q * ev) @ Q.T


@pytest.mark.parametrize("d,cond", [(1, 1.0), (2, 3.0), (5, 1e2), (16, 1e4), (64, 1e6), (100, 1e3),
                                    (128, 1e8)])
def test_condition_number_known_spectrum(orc, d, cond):
    """NormalDensity::ComputeConditionNumber (EM2/NormalDensity.cpp:67-71):
    |lambda_max / lambda_min| against the constructed spectrum, numpy's
    eigvalsh and the oracle's orc_cond.  A backward-stable solver leaves the
    smallest eigenvalue an error of about eps * lambda_max, so the relative
    tolerance grows with the condition number."""
    ev = np.geomspace(1.0, cond, d) if d > 1 else np.array([2.5])
    C = spd_with_spectrum(ev, d)
    got = fem.condition_number(C)
    tol = 1e-13 * d * max(cond, 1.0)
    assert abs(got - cond) <= tol * cond
    w = np.linalg.eigvalsh(C)
    assert abs(got - w[-1] / w[0]) <= tol * cond
    assert abs(got - orc.cond(C)) <= tol * cond


def test_condition_number_uses_lower_triangle():
    C = spd_with_spectrum(np.array([1.0, 2.0, 9.0]), 3)
    U = np.triu(C, 1)
    assert fem.condition_number(C + 5.0 * U) == fem.condition_number(C)


def test_condition_number_not_pd():
    with pytest.raises(fem.FemError) as ei:
        fem.condition_number(np.diag([1.0, -1.0, 2.0]))
    assert ei.value.code == fem.FEM_ERR_NOT_PD


def bhatt_ref(m1, C1, m2, C2):
    S = 0.5 * (C1 + C2)
    dm = m2 - m1
    d1 = 0.125 * dm @ np.linalg.solve(S, dm)
    _, l1 = np.linalg.slogdet(C1)
    _, l2 = np.linalg.slogdet(C2)
    _, ls = np.linalg.slogdet(S)
    return max(d1 + 0.5 * (ls - 0.5 * (l1 + l2)), 0.0)


@pytest.mark.parametrize("d", [1, 3, 16, 40])
def test_bhatt_dist_closed_form(d):
    """BhattDist::ComputeDist (EM2/NormalDensity.cpp:128-138) against the
    textbook Bhattacharyya distance of two Gaussians."""
    rng = np.random.default_rng(d)
    A, B = rng.standard_normal((2, d, d))
    C1 = A @ A.T / d + 0.5 * np.eye(d)
    C2 = B @ B.T / d + 0.2 * np.eye(d)
    m1, m2 = rng.standard_normal((2, d))
    got = fem.bhatt_dist(m1, C1, m2, C2)
    assert abs(got - bhatt_ref(m1, C1, m2, C2)) <= 1e-11 * max(1.0, got)
    assert abs(got - fem.bhatt_dist(m2, C2, m1, C1)) <= 1e-11 * max(1.0, got)  # symmetric


def test_bhatt_dist_special_cases():
    d = 4
    C = np.diag([1.0, 2.0, 3.0, 4.0])
    m = np.arange(d, dtype=float)
    assert fem.bhatt_dist(m, C, m, C) == 0.0  # identical densities (clamped at 0)
    # equal covariances: only the Mahalanobis term, (dm^T C^-1 dm) / 8
    dm = np.array([1.0, 0.0, 0.0, 2.0])
    np.testing.assert_allclose(fem.bhatt_dist(m, C, m + dm, C), (1.0 + 4.0 / 4.0) / 8.0, rtol=1e-14)


@pytest.mark.gpu
def test_classes_are_posterior_argmax():
    """RetrieveClasses (EM2/EMAlgorithm.cpp:107-122) on the device: the first
    component of largest posterior, in every precision form."""
    n, d, K = 30000, 6, 5
    X, w, mu, cov = fem.gengmm(n, d, K, -4.0, 4.0, 0.5, 61)
    for prec in (fem.PREC_FP64, fem.PREC_FP32, fem.PREC_BF16X3):
        print("classes precision", prec, flush=True)
        with fem.Context(n, d, K, precision=prec) as ctx:
            ctx.upload(X)
            ctx.set_params(w, mu, cov)
            ctx.factor()
            ctx.estep()
            cls = ctx.classes()
            R = ctx.posteriors()
        np.testing.assert_array_equal(cls, np.argmax(R, axis=1))
