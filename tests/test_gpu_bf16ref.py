"""Reference-centred BF16x3 E-step against the FP64 oracle (GPU).

In the BF16x3 form, E-steps after the first (for d <= 64 and component counts
whose stage fits, K <= 75 at d = 64, 118 at d = 48) centre each point once on a reference
component a (its component of largest log-density in the previous E-step)
and add s_k[a] = U_k^T (mu_a,hi - mu_k), computed in FP64, to the
accumulators.  The densities are the same function of the parameters; only
the rounding differs.  Stated tolerances are the BF16x3 form's one-step ones:
  log-likelihood <= 2e-6 relative, posteriors <= 5e-3 absolute,
  weights <= 1e-5, means <= 1e-5, covariances <= 2e-5 (relative, max-norm);
LL trajectories of both kinds of E-step agree to 2e-6 per iteration."""
import numpy as np
import pytest

import fastem_amd as fem

pytestmark = pytest.mark.gpu

TOL = {"ll": 2e-6, "R": 5e-3, "w": 1e-5, "mu": 1e-5, "cov": 2e-5}


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def start(w, mu, cov, seed):
    rng = np.random.default_rng(seed)
    K = len(w)
    return np.full(K, 1.0 / K), mu + 0.3 * rng.standard_normal(mu.shape), cov * 1.2


def balanced(n, d, K, lo, hi, scale, seed):
    """n points, n/K (rounded) from each of K Gaussians with means uniform in
    [lo, hi]^d and random full covariances (scale^2 / 4 I plus a random
    Gram matrix of the same size), in shuffled order: every component holds
    enough points for a positive definite covariance."""
    rng = np.random.default_rng(seed)
    mu = rng.uniform(lo, hi, (K, d))
    A = rng.standard_normal((K, d, d)) * (scale / np.sqrt(d))
    cov = A @ np.transpose(A, (0, 2, 1)) + (scale * scale / 4.0) * np.eye(d)
    lab = rng.permutation(np.arange(n) % K)
    L = np.linalg.cholesky(cov)
    X = mu[lab] + np.einsum("nij,nj->ni", L[lab], rng.standard_normal((n, d)))
    return X.astype(np.float32), np.bincount(lab, minlength=K) / n, mu, cov


# (n, d, K, lo, hi, scale): two reach the largest K the form takes at d = 48
# and d = 64; the last two have a single component (one reference, one seed
# row) and row counts that are not multiples of the 256-point workgroup
SHAPES = [(100000, 16, 8, -10.0, 10.0, 0.1), (50000, 64, 16, -20.0, 20.0, 0.05),
          (20000, 17, 5, -10.0, 10.0, 0.1), (50000, 32, 8, -1.5, 1.5, 0.5),
          (40000, 8, 80, -20.0, 20.0, 0.05), (20000, 3, 4, -10.0, 10.0, 0.1),
          (20000, 1, 2, -10.0, 10.0, 0.1), (30000, 64, 64, -20.0, 20.0, 0.05),
          (30000, 48, 118, -20.0, 20.0, 0.05), (30000, 64, 75, -20.0, 20.0, 0.05),
          (5000, 5, 1, -1.0, 1.0, 1.0), (12345, 64, 1, -1.0, 1.0, 1.0)]


@pytest.mark.parametrize("n,d,K,lo,hi,scale", SHAPES)
def test_reference_estep_vs_oracle(orc, n, d, K, lo, hi, scale):
    """E-step and M-step from the parameters after one EM iteration, the
    E-step taking its references from the first one."""
    print(f"bf16x3-ref one step n={n} d={d} K={K}", flush=True)
    X, w, mu, cov = balanced(n, d, K, lo, hi, scale, 91 + d + K)
    w0, mu0, cov0 = start(w, mu, cov, 2)
    with fem.Context(n, d, K, fem.PREC_BF16X3) as ctx:
        ctx.upload(X)
        ctx.set_params(w0, mu0, cov0)
        ctx.factor()
        ctx.estep()
        assert ctx.estep_form() == "bf16x3"
        w1, mu1, cov1 = ctx.mstep()
        ctx.factor()
        ll = ctx.estep()
        assert ctx.estep_form() == "bf16x3-ref"
        R = ctx.posteriors()
        w2, mu2, cov2 = ctx.mstep()
    U, ld = orc.factor_all(cov1)
    ll_ref, R_ref = orc.estep(X, w1, mu1, U, ld)
    w_ref, mu_ref, cov_ref = orc.mstep(X, R_ref)
    errs = {"ll": abs(ll - ll_ref) / abs(ll_ref), "R": float(np.max(np.abs(R - R_ref))),
            "w": rel(w2, w_ref), "mu": rel(mu2, mu_ref), "cov": rel(cov2, cov_ref)}
    print("bf16x3-ref errors", errs, flush=True)
    for key in errs:
        assert errs[key] <= TOL[key], (key, errs[key])


def test_reference_form_not_taken_where_it_does_not_fit():
    for d, K in ((64, 76), (48, 119), (80, 4)):
        print(f"bf16x3 d={d} K={K}: per-component form", flush=True)
        n = 200 * K
        X, w, mu, cov = balanced(n, d, K, -20.0, 20.0, 0.05, 5)
        with fem.Context(n, d, K, fem.PREC_BF16X3) as ctx:
            ctx.upload(X)
            ctx.set_params(w, mu, cov)
            for _ in range(2):
                ctx.factor()
                ctx.estep()
                assert ctx.estep_form() == "bf16x3"
                ctx.mstep(copy=False)


def run_trace(X, K, w0, mu0, cov0, iters):
    n, d = X.shape
    lls, forms = [], []
    with fem.Context(n, d, K, fem.PREC_BF16X3) as ctx:
        ctx.upload(X)
        ctx.set_params(w0, mu0, cov0)
        for _ in range(iters):
            ctx.factor()
            lls.append(ctx.estep())
            forms.append(ctx.estep_form())
            ctx.mstep(copy=False)
    return np.array(lls), forms


def test_reference_trajectory_matches_component_form(orc, monkeypatch):
    """Twelve EM iterations from the perturbed generating mixture
    (N=200,000, d=64, K=32, overlapping balanced components: fractional
    posteriors), once with
    reference-centred E-steps and once (FASTEM_ESTEPB_REF=0) with every pair
    centred on its own mean: per-iteration LLs within 2e-6 of each other and
    of the oracle's trajectory over the oracle's sweeps."""
    n, d, K, iters = 200000, 64, 32, 12
    print("bf16x3-ref trajectory", flush=True)
    X, w, mu, cov = balanced(n, d, K, -1.0, 1.0, 1.0, 17)
    w0, mu0, cov0 = start(w, mu, cov, 9)
    ll_ref, forms = run_trace(X, K, w0, mu0, cov0, iters)
    assert forms == ["bf16x3"] + ["bf16x3-ref"] * (iters - 1)
    monkeypatch.setenv("FASTEM_ESTEPB_REF", "0")
    ll_cmp, forms = run_trace(X, K, w0, mu0, cov0, iters)
    assert forms == ["bf16x3"] * iters
    # the oracle stops once its LL repeats exactly (about 10 sweeps here)
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, svd=False)
    m = len(ref.ll_trace)
    assert m >= 8
    e1 = np.abs(ll_ref - ll_cmp) / np.abs(ll_cmp)
    e2 = np.abs(ll_ref[:m] - ref.ll_trace) / np.abs(ref.ll_trace)
    print("per-iteration LL rel diff vs component form", e1.max(), "vs oracle", e2.max(), flush=True)
    assert e1.max() <= 2e-6 and e2.max() <= 2e-6


def test_references_reset_when_data_or_means_change():
    n, d, K = 20000, 8, 4
    X, w, mu, cov = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 3)
    with fem.Context(n, d, K, fem.PREC_BF16X3) as ctx:
        ctx.upload(X)

        def step(expect):
            ctx.factor()
            ctx.estep()
            assert ctx.estep_form() == expect
            ctx.mstep(copy=False)

        ctx.set_params(w, mu, cov)
        step("bf16x3")
        step("bf16x3-ref")
        ctx.set_params(w, mu, cov)
        step("bf16x3")
        step("bf16x3-ref")
        ctx.upload(X)
        step("bf16x3")
        step("bf16x3-ref")
        ctx.set_precision(fem.PREC_BF16X3)
        step("bf16x3")
        ctx.kmeans_init(seed=4)
        step("bf16x3")
        step("bf16x3-ref")
        ctx.set_precision(fem.PREC_FP64)
        step("fp64")


def test_reference_full_size_loglik_vs_fp64():
    """At N = 4,194,304, d = 64, K = 64 (the EM benchmark's workload): the
    reference-centred E-step's LL against the FP64 form's at the same
    parameters, and its classes against the FP64 posteriors' argmax."""
    n, d, K = 4194304, 64, 64
    print("bf16x3-ref full size", flush=True)
    X, w, mu, cov = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 2024)
    w0, mu0, cov0 = start(w, mu, cov, 4)
    with fem.Context(n, d, K, fem.PREC_BF16X3) as cb:
        cb.upload(X)
        cb.set_params(w0, mu0, cov0)
        cb.factor()
        cb.estep()
        w1, mu1, cov1 = cb.mstep()
        cb.factor()
        llb = cb.estep()
        assert cb.estep_form() == "bf16x3-ref"
        clsb = cb.classes()
    with fem.Context(n, d, K, fem.PREC_FP64) as c64:
        c64.upload(X)
        c64.set_params(w1, mu1, cov1)
        c64.factor()
        ll64 = c64.estep()
        cls64 = c64.classes()
    e = abs(llb - ll64) / abs(ll64)
    frac = float(np.mean(clsb != cls64))
    print(f"LL rel diff {e:.2e}, class disagreement {frac:.2e}", flush=True)
    assert e <= 2e-6
    assert frac <= 1e-5
