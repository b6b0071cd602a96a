"""Writes the golden fixtures that pin the FP64 oracle.

Every expected value here comes from third-party textbook implementations
(numpy, scipy, scikit-learn), not from the oracle and not from the fastem
reference (which ships no tests or data and cannot be built here).  Run from
the repository root:  python tests/golden/make_golden.py
"""
import os
import warnings

import numpy as np
from scipy.special import logsumexp
from scipy.stats import multivariate_normal

HERE = os.path.dirname(os.path.abspath(__file__))


def spd(rng, d, jitter=0.1):
    A = rng.standard_normal((d, d))
    return A @ A.T / d + jitter * np.eye(d)


def mixture_data(rng, n, d, K, spread=6.0):
    w = rng.dirichlet(np.ones(K))
    mu = rng.uniform(-spread, spread, size=(K, d))
    cov = np.stack([spd(rng, d, 0.3) for _ in range(K)])
    lab = rng.choice(K, size=n, p=w)
    X = np.empty((n, d))
    for k in range(K):
        idx = lab == k
        X[idx] = rng.multivariate_normal(mu[k], cov[k], size=idx.sum())
    return X.astype(np.float32), w, mu, cov


def factor_fixture(rng):
    out = {}
    for d in (1, 2, 3, 8, 17):
        C = spd(rng, d)
        L = np.linalg.cholesky(C)
        out[f"C_{d}"] = C
        out[f"U_{d}"] = np.linalg.inv(L).T
        out[f"logdet_{d}"] = np.linalg.slogdet(C)[1]
        out[f"eig_{d}"] = np.linalg.eigvalsh(C)
    return out


def estep_fixture(rng, n, d, K, tag):
    X, w, mu, cov = mixture_data(rng, n, d, K)
    # perturb the parameters so they are not the generating ones
    mu = mu + 0.3 * rng.standard_normal(mu.shape)
    cov = np.stack([c + 0.05 * np.eye(d) for c in cov])
    Xd = X.astype(np.float64)
    logp = np.stack([multivariate_normal(mu[k], cov[k]).logpdf(Xd).reshape(n) + np.log(w[k])
                     for k in range(K)], axis=1)
    lse = logsumexp(logp, axis=1)
    R = np.exp(logp - lse[:, None])
    return {f"{tag}_X": X, f"{tag}_w": w, f"{tag}_mu": mu, f"{tag}_cov": cov,
            f"{tag}_logp": logp, f"{tag}_LL": lse.sum(), f"{tag}_R": R}


def mstep_fixture(rng, n, d, K, tag):
    X, _, _, _ = mixture_data(rng, n, d, K)
    R = rng.dirichlet(np.ones(K), size=n)
    Xd = X.astype(np.float64)
    nk = R.sum(0)
    mu = np.stack([np.average(Xd, axis=0, weights=R[:, k]) for k in range(K)])
    cov = np.stack([np.cov(Xd, rowvar=False, aweights=R[:, k], bias=True).reshape(d, d)
                    for k in range(K)])
    return {f"{tag}_X": X, f"{tag}_R": R, f"{tag}_w": nk / n, f"{tag}_mu": mu,
            f"{tag}_cov": cov}


def sklearn_trajectory(rng, n, d, K, iters, tag):
    """scikit-learn GaussianMixture stepped one EM iteration per fit() call.
    lower_bound_ is the mean log-likelihood evaluated in the E-step before the
    M-step, the same point in the iteration at which fastem reports LogL."""
    from sklearn.mixture import GaussianMixture
    X, w, mu, cov = mixture_data(rng, n, d, K, spread=4.0)
    w0 = np.full(K, 1.0 / K)
    mu0 = X[rng.choice(n, K, replace=False)].astype(np.float64)
    cov0 = np.stack([np.cov(X.astype(np.float64), rowvar=False).reshape(d, d)] * K)
    prec0 = np.stack([np.linalg.inv(c) for c in cov0])
    gm = GaussianMixture(n_components=K, covariance_type="full", reg_covar=0.0, tol=0.0,
                         max_iter=1, warm_start=True, weights_init=w0, means_init=mu0,
                         precisions_init=prec0)
    lls, ws, mus, covs = [], [], [], []
    Xd = X.astype(np.float64)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for _ in range(iters):
            gm.fit(Xd)
            lls.append(gm.lower_bound_ * n)
            ws.append(gm.weights_.copy())
            mus.append(gm.means_.copy())
            covs.append(gm.covariances_.copy())
    return {f"{tag}_X": X, f"{tag}_w0": w0, f"{tag}_mu0": mu0, f"{tag}_cov0": cov0,
            f"{tag}_ll": np.array(lls), f"{tag}_w": np.array(ws), f"{tag}_mu": np.array(mus),
            f"{tag}_cov": np.array(covs)}


def main():
    rng = np.random.default_rng(20240611)
    fx = {}
    fx.update(factor_fixture(rng))
    fx.update(estep_fixture(rng, 300, 3, 2, "e1"))
    fx.update(estep_fixture(rng, 500, 5, 4, "e2"))
    fx.update(mstep_fixture(rng, 400, 4, 3, "m1"))
    fx.update(sklearn_trajectory(rng, 2000, 3, 3, 12, "t1"))
    fx.update(sklearn_trajectory(rng, 3000, 6, 4, 8, "t2"))
    np.savez_compressed(os.path.join(HERE, "oracle_golden.npz"), **fx)
    print("wrote", len(fx), "arrays")


if __name__ == "__main__":
    main()
