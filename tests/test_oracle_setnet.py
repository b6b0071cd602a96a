"""Pins the FP64 numpy set-network reference (oracle/setnet_ref.py) by central
finite differences on every parameter block and on the loss outputs, and by
its invariances (permutation of the points, of the target components)."""
import numpy as np
import pytest

from oracle import setnet_ref as sr


def setup(B=2, n=7, d=3, K=4, J=3, seed=0):
    rng = np.random.default_rng(seed)
    p = sr.init_params(d, K, seed, width=8)
    for k in p:
        p[k] = p[k] + 0.1 * rng.standard_normal(p[k].shape)
    X = rng.standard_normal((B, n, d))
    M = rng.standard_normal((B, J, d)) * 2
    V = rng.standard_normal((B, J, d)) * 0.3
    pw = rng.dirichlet(np.ones(J), size=B)
    return p, X, M, V, pw


def point_targets(X, M):
    """Nearest target mean of every point (the point head's target)."""
    D = ((X[:, :, None, :] - M[:, None, :, :]) ** 2).sum(-1)
    idx = D.argmin(axis=2)
    return np.take_along_axis(M, idx[:, :, None], axis=1)


def total_loss(p, X, M, V, pw):
    mu, lv, cache = sr.forward(p, X)
    L, _, _ = sr.loss_and_grad_outputs(mu, lv, M, V, pw)
    Lp, _ = sr.point_loss_and_grad(cache["m"], point_targets(X, M), gam=0.7)
    return L + Lp


def test_backward_matches_finite_differences():
    p, X, M, V, pw = setup()
    mu, lv, cache = sr.forward(p, X)
    L, dmu, dlv = sr.loss_and_grad_outputs(mu, lv, M, V, pw)
    _, dm = sr.point_loss_and_grad(cache["m"], point_targets(X, M), gam=0.7)
    gr = sr.backward(p, cache, dmu, dlv, dm)
    h = 1e-6
    for name in p:
        flat = p[name].ravel()
        for idx in range(min(flat.size, 12)):
            old = flat[idx]
            flat[idx] = old + h
            Lp = total_loss(p, X, M, V, pw)
            flat[idx] = old - h
            Lm = total_loss(p, X, M, V, pw)
            flat[idx] = old
            fd = (Lp - Lm) / (2 * h)
            an = gr[name].ravel()[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, an)


def test_output_gradient_matches_finite_differences():
    p, X, M, V, pw = setup(seed=3)
    mu, lv, _ = sr.forward(p, X)
    L, dmu, dlv = sr.loss_and_grad_outputs(mu, lv, M, V, pw)
    h = 1e-7
    for arr, grad in ((mu, dmu), (lv, dlv)):
        for idx in [(0, 0, 0), (1, 2, 1), (0, 3, 2)]:
            old = arr[idx]
            arr[idx] = old + h
            Lp, _, _ = sr.loss_and_grad_outputs(mu, lv, M, V, pw)
            arr[idx] = old - h
            Lm, _, _ = sr.loss_and_grad_outputs(mu, lv, M, V, pw)
            arr[idx] = old
            assert abs((Lp - Lm) / (2 * h) - grad[idx]) <= 1e-5


def test_point_head_is_a_convex_combination_and_equivariant():
    p, X, M, V, pw = setup(seed=4)
    _, _, cache = sr.forward(p, X)
    beta, m = cache["beta"], cache["m"]
    np.testing.assert_allclose(beta.sum(axis=2), 1.0, rtol=0, atol=1e-12)
    assert (beta >= 0).all()
    # m_i lies in the convex hull: bounded by the per-coordinate min/max of the set
    assert (m <= X.max(axis=1, keepdims=True) + 1e-12).all()
    assert (m >= X.min(axis=1, keepdims=True) - 1e-12).all()
    perm = np.random.default_rng(2).permutation(X.shape[1])
    _, _, c2 = sr.forward(p, X[:, perm])
    np.testing.assert_allclose(c2["m"], m[:, perm], rtol=1e-12, atol=1e-12)
    # with Wp = 0 every score is equal: m_i is the set mean
    p0 = dict(p)
    p0["Wp"] = np.zeros_like(p["Wp"])
    _, _, c0 = sr.forward(p0, X)
    np.testing.assert_allclose(c0["m"], np.broadcast_to(X.mean(axis=1, keepdims=True), m.shape),
                               rtol=1e-12, atol=1e-12)


def test_invariances():
    p, X, M, V, pw = setup(seed=5)
    mu, lv, _ = sr.forward(p, X)
    perm = np.random.default_rng(1).permutation(X.shape[1])
    mu2, lv2, _ = sr.forward(p, X[:, perm])
    np.testing.assert_allclose(mu, mu2, rtol=1e-12, atol=1e-12)
    L1, _, _ = sr.loss_and_grad_outputs(mu, lv, M, V, pw)
    jp = [2, 0, 1]
    L2, _, _ = sr.loss_and_grad_outputs(mu, lv, M[:, jp], V[:, jp], pw[:, jp])
    assert abs(L1 - L2) <= 1e-12 * abs(L1)


def test_flatten_roundtrip():
    p = sr.init_params(5, 6, 0)
    theta = sr.flatten(p, 5, 6)
    q = sr.unflatten(theta, 5, 6)
    for k in p:
        assert np.array_equal(p[k], q[k])


def test_round_bf16_known_values():
    """bf16 keeps 8 significant bits: 1 + 2^-8 is a tie (to even: 1), 1 +
    3 * 2^-8 rounds up to 1 + 2^-6, 1 + 2^-8 + 2^-20 rounds up to 1 + 2^-7;
    representable values pass unchanged; signs are kept."""
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.15625,
                  3.0 * 2.0 ** -130])
    want = np.array([1.0, 1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7, -2.5, 0.15625, 3.0 * 2.0 ** -130])
    np.testing.assert_array_equal(sr.round_bf16(x), want)


def test_round_bf16_is_nearest_even_on_random_floats():
    """Against the definition: the nearest of the two bf16 neighbours (FP32
    with the low 16 bits cleared, and the next one up in magnitude), the one
    with an even last bit on a tie."""
    rng = np.random.default_rng(3)
    f = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    u = f.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    x = f.astype(np.float64)
    dlo, dhi = np.abs(x - lo), np.abs(hi - x)
    even_lo = ((u >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(even_lo, lo, hi)))
    np.testing.assert_array_equal(sr.round_bf16(f), want)


def test_forward_bf16_reduces_to_forward_on_bf16_data():
    """With parameters and inputs already bf16 values, only the rounding of
    intermediate products separates forward_bf16 from forward: the outputs
    agree to bf16 resolution of those intermediates, and differ from them
    when the inputs are not bf16 values."""
    d, K = 8, 4
    p = sr.init_params(d, K, seed=2)
    for k in p:
        p[k] = sr.round_bf16(p[k])
    X = sr.round_bf16(np.random.default_rng(4).standard_normal((2, 40, d)))
    mu, lv, c = sr.forward(p, X)
    mu_b, lv_b, m_b = sr.forward_bf16(p, X)
    assert np.abs(mu_b - mu).max() <= 2e-2 * np.abs(mu).max()
    assert np.abs(m_b - c["m"]).max() <= 2e-2 * np.abs(X).max()
    assert np.abs(mu_b - mu).max() > 0.0
