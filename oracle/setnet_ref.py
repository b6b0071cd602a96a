"""FP64 numpy restatement of the learned initializer's set network, forward and
backward, and of its training loss.

TEST INFRASTRUCTURE ONLY: the checker for the HIP training kernels
(fastem_amd/csrc/fem_setnet.hip).  Pinned by central finite differences in
tests/test_oracle_setnet.py.  The network has no counterpart in the fastem
reference.  forward_bf16 restates the inference forward whose products take
operands rounded to bf16 (round_bf16, pinned on known values and against a
bit-level definition of round-to-nearest-even in the same test file).

Network (per point set of n normalised points x_i in R^d, width W = 64, K
outputs):
  h1_i = gelu(x_i W1 + b1), h2_i = gelu(h1_i W2 + b2)            per point
  g    = mean_i h2_i                                            set context
  q_k  = S_k + g Wg                                             K seed queries
  a_ik = softmax_i(q_k . h2_i / sqrt(W))                        attention over points
  z_k  = sum_i a_ik h2_i
  mu_k = z_k Wm + bm,  lv_k = z_k Wv + bv                       means, log-variances
Point head (a learned mean-shift step over the set):
  b_ij = softmax_j((h2_i Wp) . (h2_j Wq) / sqrt(W)),  m_i = sum_j b_ij x_j
Loss against a fitted mixture (means M_j, log-variances V_j, weights p_j):
  L = sum_j p_j ||mu_{k(j)} - M_j||^2  +  (1/K) sum_k min_j ||mu_k - M_j||^2
      + lam * sum_j p_j ||lv_{k(j)} - V_j||^2 / d,   k(j) = argmin_k ||mu_k - M_j||^2
      + gam * (1/n) sum_i ||m_i - T_i||^2,            T_i = M_{c(i)} (point i's cluster mean)
"""
import numpy as np

W = 64


def param_shapes(d, K, width=W):
    return [("W1", (d, width)), ("b1", (width,)), ("W2", (width, width)), ("b2", (width,)),
            ("S", (K, width)), ("Wg", (width, width)), ("Wm", (width, d)), ("bm", (d,)),
            ("Wv", (width, d)), ("bv", (d,)), ("Wp", (width, width)), ("Wq", (width, width))]


def unflatten(theta, d, K, width=W):
    out, o = {}, 0
    for name, shp in param_shapes(d, K, width):
        size = int(np.prod(shp))
        out[name] = theta[o:o + size].reshape(shp)
        o += size
    return out


def flatten(p, d, K, width=W):
    return np.concatenate([p[name].ravel() for name, _ in param_shapes(d, K, width)])


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(p, X):
    """X: (B, n, d) normalised point sets.  Returns (mu, lv, cache)."""
    width = p["W2"].shape[0]
    z1 = X @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    g = h2.mean(axis=1)                                   # (B, W)
    q = p["S"][None] + (g @ p["Wg"])[:, None, :]          # (B, K, W)
    s = np.einsum("bkw,biw->bki", q, h2) / np.sqrt(width)
    s = s - s.max(axis=2, keepdims=True)
    a = np.exp(s)
    a = a / a.sum(axis=2, keepdims=True)                  # softmax over points
    z = np.einsum("bki,biw->bkw", a, h2)
    mu = z @ p["Wm"] + p["bm"]
    lv = z @ p["Wv"] + p["bv"]
    P = h2 @ p["Wp"]
    Q = h2 @ p["Wq"]
    e = np.einsum("biw,bjw->bij", P, Q) / np.sqrt(width)
    e = e - e.max(axis=2, keepdims=True)
    bm_ = np.exp(e)
    bm_ = bm_ / bm_.sum(axis=2, keepdims=True)            # softmax over j
    m = np.einsum("bij,bjd->bid", bm_, X)
    cache = dict(X=X, z1=z1, h1=h1, z2=z2, h2=h2, g=g, q=q, a=a, z=z, P=P, Q=Q, beta=bm_, m=m)
    return mu, lv, cache


def round_bf16(x):
    """Values rounded to bfloat16 (8 significant bits, round to nearest, ties
    to even) through their FP32 bit patterns; returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, X):
    """The inference forward: every product of the per-point layers and of the
    point head (x W1, h1 W2, h2 Wp, h2 Wq, the scores P Q^T and the weighted
    sum beta X) takes operands rounded to bf16; accumulation, biases, GELU,
    pooling, the seed-query attention, the heads and the softmax are exact.
    Returns (mu, lv, m)."""
    r = round_bf16
    width = p["W2"].shape[0]
    h1 = gelu(r(X) @ r(p["W1"]) + p["b1"])
    h2 = gelu(r(h1) @ r(p["W2"]) + p["b2"])
    g = h2.mean(axis=1)
    q = p["S"][None] + (g @ p["Wg"])[:, None, :]
    s = np.einsum("bkw,biw->bki", q, h2) / np.sqrt(width)
    s = s - s.max(axis=2, keepdims=True)
    a = np.exp(s)
    a = a / a.sum(axis=2, keepdims=True)
    z = np.einsum("bki,biw->bkw", a, h2)
    mu = z @ p["Wm"] + p["bm"]
    lv = z @ p["Wv"] + p["bv"]
    P = r(h2) @ r(p["Wp"])
    Q = r(h2) @ r(p["Wq"])
    e = r(P) @ np.swapaxes(r(Q), 1, 2) / np.sqrt(width)
    e = e - e.max(axis=2, keepdims=True)
    be = np.exp(e)
    be = be / be.sum(axis=2, keepdims=True)
    m = r(be) @ r(X)
    return mu, lv, m


def point_loss_and_grad(m, T, gam=1.0):
    """gam/n * sum_i ||m_i - T_i||^2 summed over the batch, and d/dm."""
    n = m.shape[1]
    diff = m - T
    return gam * float((diff * diff).sum()) / n, 2.0 * gam * diff / n


def loss_and_grad_outputs(mu, lv, M, V, pw, lam=0.5):
    """Loss summed over the batch and its gradient w.r.t. mu and lv."""
    B, K, d = mu.shape
    L = 0.0
    dmu = np.zeros_like(mu)
    dlv = np.zeros_like(lv)
    for b in range(B):
        D = ((mu[b][:, None, :] - M[b][None, :, :]) ** 2).sum(-1)   # (K, J)
        kj = D.argmin(axis=0)                                          # best prediction per target
        jk = D.argmin(axis=1)                                          # best target per prediction
        for j in range(M.shape[1]):
            k = kj[j]
            diff = mu[b, k] - M[b, j]
            L += pw[b, j] * (diff @ diff)
            dmu[b, k] += 2.0 * pw[b, j] * diff
            dv = lv[b, k] - V[b, j]
            L += lam * pw[b, j] * (dv @ dv) / d
            dlv[b, k] += 2.0 * lam * pw[b, j] * dv / d
        for k in range(K):
            j = jk[k]
            diff = mu[b, k] - M[b, j]
            L += (diff @ diff) / K
            dmu[b, k] += 2.0 * diff / K
    return L, dmu, dlv


def backward(p, cache, dmu, dlv, dm=None):
    """Gradients of sum(dmu*mu) + sum(dlv*lv) + sum(dm*m) w.r.t. every parameter."""
    width = p["W2"].shape[0]
    X, z1, h1, z2, h2 = cache["X"], cache["z1"], cache["h1"], cache["z2"], cache["h2"]
    g, q, a, z = cache["g"], cache["q"], cache["a"], cache["z"]
    n = X.shape[1]
    if dm is None:
        dm = np.zeros_like(cache["m"])
    gr = {}
    gr["Wm"] = np.einsum("bkw,bkd->wd", z, dmu)
    gr["bm"] = dmu.sum(axis=(0, 1))
    gr["Wv"] = np.einsum("bkw,bkd->wd", z, dlv)
    gr["bv"] = dlv.sum(axis=(0, 1))
    dz = dmu @ p["Wm"].T + dlv @ p["Wv"].T                        # (B, K, W)
    dh2 = np.einsum("bki,bkw->biw", a, dz)                         # through z = a h2
    da = np.einsum("bkw,biw->bki", dz, h2)
    ds = a * (da - (a * da).sum(axis=2, keepdims=True))           # softmax backward
    ds = ds / np.sqrt(width)
    dq = np.einsum("bki,biw->bkw", ds, h2)
    dh2 += np.einsum("bki,bkw->biw", ds, q)
    gr["S"] = dq.sum(axis=0)
    dqs = dq.sum(axis=1)                                           # (B, W)
    gr["Wg"] = g.T @ dqs
    dg = dqs @ p["Wg"].T
    dh2 += dg[:, None, :] / n
    # point head
    P, Q, beta = cache["P"], cache["Q"], cache["beta"]
    db = np.einsum("bid,bjd->bij", dm, X)
    de = beta * (db - (beta * db).sum(axis=2, keepdims=True)) / np.sqrt(width)
    dP = np.einsum("bij,bjw->biw", de, Q)
    dQ = np.einsum("bij,biw->bjw", de, P)
    gr["Wp"] = np.einsum("biw,biv->wv", h2, dP)
    gr["Wq"] = np.einsum("biw,biv->wv", h2, dQ)
    dh2 += dP @ p["Wp"].T + dQ @ p["Wq"].T
    dz2 = dh2 * gelu_grad(z2)
    gr["W2"] = np.einsum("biw,biv->wv", h1, dz2)
    gr["b2"] = dz2.sum(axis=(0, 1))
    dh1 = dz2 @ p["W2"].T
    dz1 = dh1 * gelu_grad(z1)
    gr["W1"] = np.einsum("bid,biw->dw", X, dz1)
    gr["b1"] = dz1.sum(axis=(0, 1))
    return gr


def init_params(d, K, seed, width=W):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in param_shapes(d, K, width):
        if name.startswith("b"):
            p[name] = np.zeros(shp)
        elif name == "S":
            p[name] = rng.standard_normal(shp)
        else:
            p[name] = rng.standard_normal(shp) / np.sqrt(shp[0])
    return p
