"""Pins the FP64 oracle (oracle/oracle.c) against golden vectors from numpy,
scipy and scikit-learn, the C++ standard's known answers, closed forms and
brute force.  CPU only."""
import numpy as np
import pytest


# ---- Util/Rand.cpp: std::mt19937_64 ------------------------------------------------

def test_mt19937_64_known_answer(orc):
    # C++11 [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937_64 (seed 5489) produces 9981545732273789042.
    r = orc.Rand(5489)
    for _ in range(9999):
        r.next_u64()
    assert r.next_u64() == 9981545732273789042


def test_canonical_mapping(orc):
    # generate_canonical<double,53> over a 64-bit engine: double(x) * 2^-64.
    a, b = orc.Rand(7), orc.Rand(7)
    for _ in range(1000):
        x = a.next_u64()
        u = b.canonical()
        assert u == float(x) / 2.0 ** 64
        assert 0.0 <= u < 1.0


# ---- Util/OpenMP.h / EM2/BlockManager.cpp -----------------------------------------

@pytest.mark.parametrize("n,size", [(10, 3), (7, 7), (5, 8), (100000, 16), (1, 4)])
def test_find_items_partition(orc, n, size):
    spans = [orc.find_items(n, size, i) for i in range(size)]
    pos = 0
    for off, cnt in spans:
        assert off == pos
        pos += cnt
    assert pos == n
    cnts = [c for _, c in spans]
    assert max(cnts) - min(cnts) <= 1
    assert cnts == sorted(cnts, reverse=True)


def test_block_params(orc):
    assert orc.parse_block_params(None, 1000, 4, 2) == (1, 1)
    assert orc.parse_block_params("3:5", 1000, 4, 2) == (3, 5)
    # nDens = floor(N*16(d+K)/(L3 MB * 2^20) * 1.2), at least 1
    n, d, K = 4194304, 64, 64
    exp = int(n * (16 * d + 16 * K) / (18 * 2 ** 20) * 1.2)
    assert orc.parse_block_params("automax:18", n, d, K) == (exp, exp)
    assert orc.parse_block_params("automax:1000", 100, 2, 2) == (1, 1)
    with pytest.raises(ValueError):
        orc.parse_block_params("bogus", 10, 2, 2)


# ---- EM2/NormalDensity.cpp --------------------------------------------------------

@pytest.mark.parametrize("d", [1, 2, 3, 8, 17])
def test_factor_vs_numpy(orc, golden, d):
    U, ld = orc.factor(golden[f"C_{d}"])
    np.testing.assert_allclose(U, golden[f"U_{d}"], rtol=1e-12, atol=1e-13)
    assert abs(ld - golden[f"logdet_{d}"]) <= 1e-12 * max(1.0, abs(ld))


def test_factor_closed_forms(orc):
    D = np.diag([4.0, 9.0, 0.25])
    U, ld = orc.factor(D)
    np.testing.assert_allclose(U, np.diag([0.5, 1 / 3.0, 2.0]), rtol=1e-15)
    assert abs(ld - np.log(4 * 9 * 0.25)) < 1e-15
    # 2x2: C = [[a, b], [b, c]], L = [[sqrt a, 0], [b/sqrt a, sqrt(c - b^2/a)]]
    a, b, c = 4.0, 2.0, 5.0
    U, ld = orc.factor(np.array([[a, b], [b, c]]))
    l11, l21 = np.sqrt(a), b / np.sqrt(a)
    l22 = np.sqrt(c - b * b / a)
    Linv = np.array([[1 / l11, 0], [-l21 / (l11 * l22), 1 / l22]])
    np.testing.assert_allclose(U, Linv.T, rtol=1e-15)
    assert abs(ld - np.log(a * c - b * b)) < 1e-15
    # U^T C U = I for a matrix built as L L^T from a known L
    rng = np.random.default_rng(3)
    L = np.tril(rng.standard_normal((6, 6)))
    np.fill_diagonal(L, np.abs(np.diag(L)) + 1.0)
    C = L @ L.T
    U, ld = orc.factor(C)
    np.testing.assert_allclose(U.T @ C @ U, np.eye(6), atol=1e-12)
    assert abs(ld - 2 * np.log(np.diag(L)).sum()) < 1e-12


def test_factor_uses_lower_triangle(orc):
    C = np.array([[2.0, 99.0], [0.5, 3.0]])  # upper garbage is ignored
    U1, l1 = orc.factor(C)
    U2, l2 = orc.factor(np.array([[2.0, 0.5], [0.5, 3.0]]))
    assert np.array_equal(U1, U2) and l1 == l2


def test_factor_not_pd_raises(orc):
    with pytest.raises(np.linalg.LinAlgError):
        orc.factor(np.array([[1.0, 2.0], [2.0, 1.0]]))
    with pytest.raises(np.linalg.LinAlgError):
        orc.factor(np.array([[np.nan, 0.0], [0.0, 1.0]]))


@pytest.mark.parametrize("d", [1, 2, 3, 8, 17])
def test_eigvals_and_cond(orc, golden, d):
    ev = orc.sym_eigvals(golden[f"C_{d}"])
    np.testing.assert_allclose(ev, golden[f"eig_{d}"], rtol=1e-10, atol=1e-12)
    e = golden[f"eig_{d}"]
    assert abs(orc.cond(golden[f"C_{d}"]) - abs(e[-1] / e[0])) <= 1e-9 * abs(e[-1] / e[0])


def test_eigvals_larger(orc):
    rng = np.random.default_rng(11)
    for d in (32, 64, 128):
        A = rng.standard_normal((d, d))
        C = A @ A.T / d + 0.05 * np.eye(d)
        np.testing.assert_allclose(orc.sym_eigvals(C), np.linalg.eigvalsh(C), rtol=1e-9,
                                   atol=1e-11)


# ---- EM2/EMAlgorithm.cpp:237-244 --------------------------------------------------

def test_regularize_closed_form(orc):
    rng = np.random.default_rng(1)
    C = np.stack([np.cov(rng.standard_normal((4, 20))) for _ in range(3)])
    g = 0.25
    out = orc.regularize(C, g)
    for k in range(3):
        exp = C[k] * (1 - g) + g * np.trace(C[k]) / 4 * np.eye(4)
        np.testing.assert_allclose(out[k], exp, rtol=1e-15, atol=1e-16)
        # the trace is preserved
        assert abs(np.trace(out[k]) - np.trace(C[k])) < 1e-13
    assert np.array_equal(orc.regularize(C, 0.0), C)


# ---- E-step: EM2/MatrixEM.cpp:39-87 ----------------------------------------------

@pytest.mark.parametrize("tag", ["e1", "e2"])
@pytest.mark.parametrize("nvt,nblk", [(1, 1), (16, 1), (3, 4)])
def test_estep_vs_scipy(orc, golden, tag, nvt, nblk):
    X, w, mu, cov = (golden[f"{tag}_{s}"] for s in ("X", "w", "mu", "cov"))
    U, ld = orc.factor_all(cov)
    ll, R = orc.estep(X, w, mu, U, ld, nvt=nvt, nblk=nblk)
    assert abs(ll - golden[f"{tag}_LL"]) <= 1e-12 * abs(ll)
    np.testing.assert_allclose(R, golden[f"{tag}_R"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(R.sum(1), 1.0, rtol=1e-14)


def test_estep_single_component_closed_form(orc):
    rng = np.random.default_rng(5)
    d = 4
    X = rng.standard_normal((200, d)).astype(np.float32)
    mu = rng.standard_normal((1, d))
    A = rng.standard_normal((d, d))
    C = A @ A.T + np.eye(d)
    U, ld = orc.factor_all(C[None])
    ll, R = orc.estep(X, np.array([1.0]), mu, U, ld)
    diff = X.astype(np.float64) - mu[0]
    maha = np.einsum("ij,jk,ik->i", diff, np.linalg.inv(C), diff)
    exp = (-0.5 * maha - 0.5 * d * np.log(2 * np.pi) - 0.5 * np.linalg.slogdet(C)[1]).sum()
    assert abs(ll - exp) <= 1e-12 * abs(exp)
    assert np.all(R == 1.0)


# ---- M-step: EM2/MatrixEM.cpp:309-378 --------------------------------------------

@pytest.mark.parametrize("nvt,nblk,log2", [(1, 1, False), (16, 1, True), (5, 3, False)])
def test_mstep_vs_numpy(orc, golden, nvt, nblk, log2):
    X, R = golden["m1_X"], golden["m1_R"]
    w, mu, cov = orc.mstep(X, R, nvt=nvt, nblk=nblk, log2red=log2)
    np.testing.assert_allclose(w, golden["m1_w"], rtol=1e-13)
    np.testing.assert_allclose(mu, golden["m1_mu"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(cov, golden["m1_cov"], rtol=1e-11, atol=1e-13)
    for k in range(len(w)):
        assert np.array_equal(cov[k], cov[k].T)  # mirrored from the upper triangle
    assert abs(w.sum() - 1.0) < 1e-14


def test_mstep_single_component_closed_form(orc):
    rng = np.random.default_rng(9)
    X = (rng.standard_normal((500, 3)) * [1, 2, 3] + [5, -1, 0]).astype(np.float32)
    w, mu, cov = orc.mstep(X, np.ones((500, 1)))
    Xd = X.astype(np.float64)
    np.testing.assert_allclose(mu[0], Xd.mean(0), rtol=1e-13)
    np.testing.assert_allclose(cov[0], np.cov(Xd, rowvar=False, bias=True), rtol=1e-12)
    assert w[0] == 1.0


def test_mstep_ntotal_scales_weights(orc, golden):
    X, R = golden["m1_X"], golden["m1_R"]
    w1, _, _ = orc.mstep(X, R)
    w2, _, _ = orc.mstep(X, R, ntotal=2 * len(X))
    np.testing.assert_allclose(w2 * 2, w1, rtol=1e-15)


# ---- Train loop: EM2/EMAlgorithm.cpp:133-191 vs scikit-learn ----------------------

@pytest.mark.parametrize("tag", ["t1", "t2"])
def test_trajectory_vs_sklearn(orc, golden, tag):
    X = golden[f"{tag}_X"]
    ll_ref = golden[f"{tag}_ll"]
    n_it = len(ll_ref)
    res = orc.train(X, golden[f"{tag}_w0"], golden[f"{tag}_mu0"], golden[f"{tag}_cov0"],
                    maxsweeps=n_it, eps=1e-300, svd=False)
    assert res.rc == n_it + 1
    np.testing.assert_allclose(res.ll_trace, ll_ref, rtol=1e-10)
    np.testing.assert_allclose(res.w, golden[f"{tag}_w"][-1], rtol=1e-9)
    np.testing.assert_allclose(res.mu, golden[f"{tag}_mu"][-1], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(res.cov, golden[f"{tag}_cov"][-1], rtol=1e-9, atol=1e-10)
    # EM never decreases the likelihood
    assert np.all(np.diff(res.ll_trace) >= -1e-9 * np.abs(res.ll_trace[1:]))


def test_train_stops_and_counts_like_reference(orc, golden):
    X = golden["t1_X"]
    a = (golden["t1_w0"], golden["t1_mu0"], golden["t1_cov0"])
    # maxsweeps exhausted: IterCounter = maxsweeps + 1 (printed sweeps = maxsweeps)
    r = orc.train(X, *a, maxsweeps=3, eps=1e-300, svd=False)
    assert r.rc == 4 and len(r.ll_trace) == 3
    # convergence break increments IterCounter once more
    r = orc.train(X, *a, maxsweeps=500, eps=1e-3, svd=False)
    assert 1 < r.rc <= 501
    tr = r.ll_trace
    e = np.abs((tr[-1] - tr[-2]) / tr[-1])
    assert e < 1e-3 or tr[-1] < tr[-2]
    # zero sweeps: nothing runs, LL stays at -DBL_MAX
    r = orc.train(X, *a, maxsweeps=0, svd=False)
    assert r.rc == 1 and r.ll == -np.finfo(np.float64).max


def test_train_failure_codes(orc, golden):
    X = golden["t1_X"]
    w0, mu0, cov0 = golden["t1_w0"], golden["t1_mu0"], golden["t1_cov0"].copy()
    bad = cov0.copy()
    bad[1] = np.array([[1.0, 2.0, 0], [2.0, 1.0, 0], [0, 0, 1.0]])
    assert orc.train(X, w0, mu0, bad, maxsweeps=5, svd=False).rc == -1
    aniso = cov0.copy()
    aniso[0] = np.diag([1.0, 1.0, 1e-4])
    assert orc.train(X, w0, mu0, aniso, maxsweeps=5, athr=1.5, svd=True).rc == -2
    # with the condition check off the same init runs
    assert orc.train(X, w0, mu0, aniso, maxsweeps=2, eps=1e-300, svd=False).rc == 3


# ---- k-means: Clust/* ------------------------------------------------------------

def test_classify_brute_force_and_ties(orc):
    rng = np.random.default_rng(2)
    X = rng.integers(-3, 4, size=(400, 3)).astype(np.float32)
    C = rng.integers(-3, 4, size=(6, 3)).astype(np.float64)
    C[3] = C[1]  # exact duplicate centroid: ties must go to the lower index
    cls = orc.classify(X, C)
    D = ((X.astype(np.float64)[:, None, :] - C[None]) ** 2).sum(-1)
    assert np.array_equal(cls, D.argmin(1))  # numpy argmin returns the first minimum
    assert not np.any(cls == 3)


def test_kmeanspp_selection_rule(orc):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((300, 2)).astype(np.float32)
    K = 5
    C = orc.kmeanspp(X, K, orc.Rand(17))
    # replay with an independent numpy implementation of the D^2 rule
    r = orc.Rand(17)
    Xd = X.astype(np.float64)
    i0 = int(r.canonical() * len(X))
    ref = [Xd[i0]]
    dmin = np.full(len(X), np.inf)
    for _ in range(1, K):
        dmin = np.minimum(dmin, ((ref[-1] - Xd) ** 2).sum(1))
        agg = np.cumsum(dmin) / dmin.sum()
        u = r.canonical()
        ref.append(Xd[np.argmax(u <= agg)])
    np.testing.assert_allclose(C, np.array(ref), rtol=1e-15)


def test_lloyd_vs_numpy(orc):
    rng = np.random.default_rng(6)
    centers = np.array([[0, 0], [10, 0], [0, 10], [10, 10]], dtype=np.float64)
    X = (centers[rng.integers(0, 4, 800)] + rng.standard_normal((800, 2))).astype(np.float32)
    C0 = X[[0, 1, 2, 3]].astype(np.float64)
    C, it, fit = orc.lloyd(X, C0, orc.Rand(1), minrel=1e-4, maxiter=30)
    Cr = C0.copy()
    Xd = X.astype(np.float64)
    best = np.finfo(np.float64).max
    for i in range(1, 100):
        lab = ((Xd[:, None] - Cr[None]) ** 2).sum(-1).argmin(1)
        f = ((Xd - Cr[lab]) ** 2).sum(1).sum() / len(X)
        Cr = np.stack([Xd[lab == k].mean(0) for k in range(4)])
        rel = (best - f) / best
        if f < best:
            best = f
        else:
            break
        if rel < 1e-4 or i == 29:
            break
    assert it == i
    np.testing.assert_allclose(C, Cr, rtol=1e-12)
    assert abs(fit - best) <= 1e-12 * best


def test_gmm_from_clusters(orc):
    rng = np.random.default_rng(8)
    X = rng.standard_normal((300, 3)).astype(np.float32)
    C = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    cls = orc.classify(X, C)
    w, mu, cov = orc.gmm_from_clusters(X, C, cls)
    Xd = X.astype(np.float64)
    for k in range(2):
        sel = Xd[cls == k]
        diff = sel - C[k]
        np.testing.assert_allclose(cov[k], diff.T @ diff / (len(sel) - 1), rtol=1e-12)
        assert w[k] == len(sel) / 300
        assert np.array_equal(mu[k], C[k])
