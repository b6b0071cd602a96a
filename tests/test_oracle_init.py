"""Pins the oracle's initializers (CPU): Forgy, Jain, the data covariance
trace, k-means|| and greedy k-means++, each against an independent numpy replay of the
reference's Rand() stream (Util/Rand.cpp:7-14, whose mt19937_64 and canonical
mapping test_oracle.py pins with the C++ standard's known answer)."""
import numpy as np
import pytest


def draws(orc, seed, count):
    r = orc.Rand(seed)
    return [r.canonical() for _ in range(count)]


def data(n, d, seed):
    return np.random.default_rng(seed).standard_normal((n, d)).astype(np.float32) * 3.0


@pytest.mark.parametrize("n,d,K,seed", [(1000, 3, 5, 1), (17, 2, 17, 4), (50000, 8, 64, 9)])
def test_forgy_replay(orc, n, d, K, seed):
    """Clust/KMeansInitializer.cpp:33-49: component k is row (long)(u_k N) of
    the k-th draw, repeats allowed; exactly K draws consumed."""
    X = data(n, d, seed)
    rng = orc.Rand(seed)
    C = orc.forgy(X, K, rng)
    u = draws(orc, seed, K + 1)
    idx = [int(v * n) for v in u[:K]]
    np.testing.assert_array_equal(C, X[idx].astype(np.float64))
    assert rng.canonical() == u[K]


def test_forgy_repeats_rows(orc):
    """With N < K the pigeonhole principle forces a repeated row, which Forgy
    keeps (no duplicate check)."""
    X = data(3, 2, 0)
    C = orc.forgy(X, 5, orc.Rand(7))
    assert len({tuple(r) for r in C}) <= 3


@pytest.mark.parametrize("n,d", [(2, 1), (1000, 4), (12345, 17)])
def test_data_trace(orc, n, d):
    X = data(n, d, n)
    ref = np.trace(np.atleast_2d(np.cov(X.astype(np.float64), rowvar=False, ddof=1)))
    assert abs(orc.data_trace(X) - ref) <= 1e-12 * abs(ref)


def jain_replay(X, K, u):
    """EM2/EMInitializer.cpp:50-76 replayed: the kept index is the first draw;
    the do/while re-draws into a shadowed variable until that value is not a
    kept index of an earlier component."""
    n = X.shape[0]
    it = iter(u)
    kept = []
    for i in range(K):
        idx = int(next(it) * n)
        while True:
            idx2 = int(next(it) * n)
            if idx2 not in kept[:i]:
                break
        kept.append(idx)
    return kept, len(u) - sum(1 for _ in it)


@pytest.mark.parametrize("n,d,K,seed", [(1000, 3, 5, 1), (6, 2, 5, 3), (20000, 16, 64, 8)])
def test_jain_replay(orc, n, d, K, seed):
    X = data(n, d, seed)
    rng = orc.Rand(seed)
    w, mu, cov = orc.jain_init(X, K, rng)
    u = draws(orc, seed, 4000)
    kept, used = jain_replay(X, K, u)
    np.testing.assert_array_equal(mu, X[kept].astype(np.float64))
    assert rng.canonical() == u[used]  # the re-draws were consumed
    np.testing.assert_array_equal(w, np.full(K, 1.0 / K))
    tr = np.trace(np.cov(X.astype(np.float64), rowvar=False, ddof=1))
    for k in range(K):
        assert np.count_nonzero(cov[k] - np.diag(np.diag(cov[k]))) == 0
        np.testing.assert_allclose(np.diag(cov[k]), tr / d / 10.0, rtol=1e-12)


def test_jain_small_n_redraws(orc):
    """N=6, K=5 makes duplicate re-draws frequent: over several seeds some run
    uses more than 2K draws, and every run consumes what the replay does."""
    X = data(6, 2, 3)
    most = 0
    for seed in range(20):
        rng = orc.Rand(seed)
        orc.jain_init(X, 5, rng)
        u = draws(orc, seed, 200)
        _, used = jain_replay(X, 5, u)
        assert rng.canonical() == u[used]
        most = max(most, used)
    assert most > 10


def sqdist_ff(A, x):
    """Clust/CentroidVector.h:79-87: float differences and squares, double
    accumulation in ascending feature order (rows of A against one row x)."""
    s = np.zeros(A.shape[0])
    for j in range(A.shape[1]):
        t = (x[j] - A[:, j]).astype(np.float32)
        s = s + (t * t).astype(np.float32).astype(np.float64)
    return s


def sqdist_df(c, A):
    """Clust/CentroidVector.h:105-116: double centroid c minus float rows A."""
    s = np.zeros(A.shape[0])
    for j in range(A.shape[1]):
        t = c[j] - A[:, j].astype(np.float64)
        s = s + t * t
    return s


def kmeansoror_replay(X, K, u, os=1.4):
    """Clust/KMeansOrOrInitializer.cpp:143-174 + PlusPlusInitializer.cpp:23-84
    replayed in numpy, one draw per row per round in row order."""
    n, d = X.shape
    it = iter(u)
    rounds = 10 if os <= 0.1 else 5
    first = int(next(it) * n)
    rows = [first]
    cd = sqdist_ff(X, X[first])          # distance of every row to the first center
    cc = np.zeros(n, dtype=np.int64)
    cost = np.cumsum(cd)[-1]             # sequential sum
    for _ in range(rounds):
        r = np.array([next(it) for _ in range(n)])
        prob = os * float(K) * cd / cost
        new = list(np.nonzero(r < prob)[0])
        base = len(rows)
        for j, i in enumerate(new):
            s = sqdist_ff(X, X[i])
            better = s < cd
            cd = np.where(better, s, cd)
            cc = np.where(better, base + j, cc)
        rows += new
        cost = np.cumsum(cd)[-1]
    m = len(rows)
    wt = np.bincount(cc, minlength=m).astype(np.float64)
    Cand = X[rows]
    C = np.zeros((K, d))
    C[0] = Cand[int(next(it) * m)]
    cd2 = np.full(m, np.inf)
    for c in range(1, K):
        cd2 = np.minimum(cd2, sqdist_df(C[c - 1], Cand))
        agg = np.cumsum(cd2 * wt)
        agg = agg / agg[-1]
        rr = next(it)
        pick = np.nonzero(rr <= agg)[0]
        C[c] = Cand[pick[0]]
    return C, m


@pytest.mark.parametrize("n,d,K,seed", [(2000, 3, 5, 2), (3000, 7, 12, 5), (1500, 2, 1, 1)])
def test_kmeansoror_replay(orc, n, d, K, seed):
    X = data(n, d, seed)
    rng = orc.Rand(seed)
    C, m = orc.kmeansoror(X, K, rng)
    u = draws(orc, seed, 1 + 5 * n + K + 5)
    C_ref, m_ref = kmeansoror_replay(X, K, u)
    assert m == m_ref
    np.testing.assert_array_equal(C, C_ref)
    # draws used: 1 + 5n + 1 (first weighted seed) + K-1
    assert rng.canonical() == u[1 + 5 * n + K]


def test_kmeansoror_candidate_count(orc):
    """Each round adds about os*K centers in expectation (the probabilities
    sum to os*K), so m is near 1 + 5*1.4*K."""
    X = data(20000, 4, 11)
    ms = [orc.kmeansoror(X, 16, orc.Rand(s))[1] for s in range(6)]
    assert 0.6 * (1 + 5 * 1.4 * 16) < np.mean(ms) < 1.4 * (1 + 5 * 1.4 * 16)


def test_kmeans_init_composite(orc):
    """KMeansEMInitializer: Forgy seeds, then the Lloyd + cluster tail on the
    same stream; weights are the cluster counts over N."""
    X = data(5000, 3, 21)
    w, mu, cov = orc.kmeans_init(X, 4, seed=3)
    rng = orc.Rand(3)
    C = orc.forgy(X, 4, rng)
    C, _, _ = orc.lloyd(X, C, rng)
    np.testing.assert_array_equal(mu, C)
    cls = orc.classify(X, C)
    np.testing.assert_array_equal(w * 5000, np.bincount(cls, minlength=4))


def _gmm_rows(n, d, K, seed):
    rng = np.random.default_rng(seed)
    mu = rng.uniform(-10, 10, (K, d))
    lab = rng.integers(0, K, n)
    return (mu[lab] + rng.standard_normal((n, d))).astype(np.float32)


@pytest.mark.parametrize("n,d,K,seed", [(3000, 4, 7, 1), (5001, 9, 12, 2)])
def test_greedy_kmeanspp_one_trial_is_kmeanspp(orc, n, d, K, seed):
    """With one candidate per seed the greedy replay is plain k-means++: the
    same seeds as the C oracle's orc_kmeanspp on the same Rand() stream."""
    X = _gmm_rows(n, d, K, seed)
    np.testing.assert_array_equal(orc.kmeanspp_greedy(X, K, orc.Rand(seed), L=1),
                                  orc.kmeanspp(X, K, orc.Rand(seed)))


def test_greedy_kmeanspp_keeps_least_potential(orc):
    """Brute force on a small set: replaying the draws independently (numpy
    cumsum rule, every candidate's potential by direct summation), each
    greedy seed is the candidate of least potential."""
    n, d, K, L, seed = 400, 3, 6, 4, 5
    X = _gmm_rows(n, d, K, seed)
    C = orc.kmeanspp_greedy(X, K, orc.Rand(seed), L=L)
    X64 = X.astype(np.float64)
    rng = orc.Rand(seed)
    first = int(rng.canonical() * n)
    np.testing.assert_array_equal(C[0], X64[first])
    for k in range(1, K):
        D = np.min([((X64 - C[j]) ** 2).sum(1) for j in range(k)], axis=0)
        agg = np.cumsum(D)
        cands = [int(np.argmax(rng.canonical() <= agg / agg[-1])) for _ in range(L)]
        pots = [np.minimum(D, ((X64 - X64[i]) ** 2).sum(1)).sum() for i in cands]
        np.testing.assert_array_equal(C[k], X64[cands[int(np.argmin(pots))]])
    # and the greedy potential is not above plain k-means++'s from the same stream
    Cp = orc.kmeanspp(X, K, orc.Rand(seed))
    pot = lambda S: np.min([((X64 - c) ** 2).sum(1) for c in S], axis=0).sum()
    assert pot(C) <= pot(Cp) * (1 + 1e-12)
