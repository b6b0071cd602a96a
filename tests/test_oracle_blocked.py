"""The blocked BLAS restatement of SimpleMatrixEM (oracle/blocked_em.py, the
CPU baseline's algorithm) against the FP64 oracle (CPU)."""
import numpy as np
import pytest

import fastem_amd as fem
from oracle import blocked_em as be


@pytest.mark.parametrize("n,size", [(10, 3), (100003, 16), (7, 8)])
def test_partition_matches_oracle(orc, n, size):
    for i in range(size):
        assert be.find_items(n, size, i) == orc.find_items(n, size, i)


@pytest.mark.parametrize("n,d,K,l3", [(4194304, 64, 64, 32.0), (100000, 16, 8, 1.0), (1000, 4, 3, 18.0)])
def test_automax_matches_oracle(orc, n, d, K, l3):
    assert be.automax_blocks(n, d, K, l3) == orc.parse_block_params(f"automax:{l3:g}", n, d, K)[0]


def test_blocked_iteration_matches_oracle(orc):
    """One EM iteration: LL, posteriors, weights, means, covariances within
    1e-12 of the oracle (only summation order differs)."""
    n, d, K = 20000, 16, 8
    X, w, mu, cov = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 5)
    w0, mu0, cov0 = np.full(K, 1.0 / K), mu + 0.3, cov * 1.2
    U, ld = orc.factor_all(cov0)
    ll_ref, R_ref = orc.estep(X, w0, mu0, U, ld)
    w_ref, mu_ref, cov_ref = orc.mstep(X, R_ref)
    with be.BlockedEM(X, K, threads=2, nblocks=3) as em:
        ll, w1, mu1, cov1 = em.iterate(w0, mu0, cov0)
        R = em.posteriors()
    assert abs(ll - ll_ref) <= 1e-12 * abs(ll_ref)
    np.testing.assert_allclose(R, R_ref, rtol=1e-11, atol=1e-14)
    for a, b in ((w1, w_ref), (mu1, mu_ref), (cov1, cov_ref)):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
