"""The learned initializer behind the C ABI and the drivers (GPU).

* fem_setnet_propose_means against an exact numpy replay of its steps on the
  oracle's Rand() stream (subsample indices, normalisation, the point head's
  BF16 forward, k-means++ seeding over the moved points);
* fem_setnet_propose against that replay followed by the oracle's Lloyd
  sweeps and cluster mixture (counts exact, covariances 1e-10);
* em2 / eminit with -I learned:<file> and -I learnedbest:<file> reproduce the
  Python (SetNet) proposal and judgement on the same seed;
* on mixtures of the learned initializer's workload (N=65,536, d=32, K=16) EM
  from -I learned needs fewer iterations than from -I kmeanspp."""
import os
import re
import subprocess

import numpy as np
import pytest

import fastem_amd as fem
from fastem_amd import formats

pytestmark = pytest.mark.gpu

NET = os.path.join(os.path.dirname(__file__), "..", "fastem_amd", "data", "setnet_d32_k16.bin")
N, D, K = 65536, 32, 16
GEN = (-10.0, 10.0, 0.1)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def net():
    from fastem_amd.setnet import SetNet
    with SetNet.load(NET) as n:
        yield n


def replay_means(net, X, n_sub, rng):
    """The proposal's steps restated with numpy on the same Rand() stream."""
    n, d = X.shape
    idx = [int(rng.canonical() * n) for _ in range(n_sub)]
    Xs = X[idx].astype(np.float64)
    cen = np.cumsum(Xs, axis=0)[-1] / n_sub          # sequential column sums
    dev = Xs - cen
    ss = np.cumsum((dev * dev).ravel())[-1]          # rows outer, columns inner
    sc = np.sqrt(ss / (n_sub * d))
    Z = (dev / sc).astype(np.float32)
    _, _, m = net.forward(Z[None], points=True, bf16=True)
    P = m[0].astype(np.float64)
    C = np.zeros((net.K, d))
    C[0] = P[int(rng.canonical() * n_sub)]
    cd = np.full(n_sub, np.inf)
    for k in range(1, net.K):
        s = np.zeros(n_sub)
        for j in range(d):
            t = C[k - 1, j] - P[:, j]
            s = s + t * t
        cd = np.where(cd > s, s, cd)
        agg = np.cumsum(cd)
        agg = agg / agg[-1]
        r = rng.canonical()
        hit = np.nonzero(r <= agg)[0]
        if len(hit):
            C[k] = P[hit[0]]
    return cen + sc * C


@pytest.mark.parametrize("seed", [3, 20_000_500])
def test_propose_means_replay(orc, net, seed):
    say(f"learned propose_means replay seed={seed}")
    X = fem.gengmm(N, D, K, *GEN, seed)[0]
    with fem.Context(N, D, K) as ctx:
        ctx.upload(X)
        C0 = net.propose_means(ctx, n_sub=2048, seed=seed)
    np.testing.assert_array_equal(C0, replay_means(net, X, 2048, orc.Rand(seed)))


def test_propose_full_pipeline_vs_oracle(orc, net):
    """Proposal, then Lloyd and the cluster mixture on one stream: equal to
    the replayed proposal followed by the oracle's Lloyd (MinRel 1e-4, at most
    29 sweeps), classification and mixture."""
    say("learned full pipeline vs oracle")
    seed = 20_000_501
    X = fem.gengmm(N, D, K, *GEN, seed)[0]
    with fem.Context(N, D, K) as ctx:
        ctx.upload(X)
        w, mu, cov, used = net.init_context(ctx, n_sub=2048, seed=seed)
    assert used == "learned"
    rng = orc.Rand(seed)
    C0 = replay_means(net, X, 2048, rng)
    C, _, _ = orc.lloyd(X, C0, rng)
    w_ref, mu_ref, cov_ref = orc.gmm_from_clusters(X, C, orc.classify(X, C))
    np.testing.assert_array_equal(w * N, w_ref * N)
    np.testing.assert_allclose(mu, mu_ref, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(cov - cov_ref)) <= 1e-10 * np.max(np.abs(cov_ref))


def run_tool(path, args, timeout=300):
    p = subprocess.run([path] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def test_em2_and_eminit_reproduce_python(net, tmp_path):
    say("em2/eminit learned vs python")
    seed = 20_000_502
    X = fem.gengmm(N, D, K, *GEN, seed)[0]
    formats.write_bin(tmp_path / "x.bin", X)
    with fem.Context(N, D, K) as ctx:
        ctx.upload(X)
        ref = net.init_context(ctx, n_sub=2048, seed=seed)
        refj = net.judge(ctx, n_sub=2048, seed=seed)
    rc, out = run_tool(fem.EM2_PATH, ["-c", K, "-r", seed, "-I", f"learned:{NET}", "-m", 0, "-S", "-o",
                                      tmp_path / "a.cmp", tmp_path / "x.bin"])
    assert rc == 0 and "Learned EM initializer" in out, out
    for a, b in zip(formats.read_cmp(tmp_path / "a.cmp"), ref[:3]):
        np.testing.assert_array_equal(a, b)
    rc, out = run_tool(fem.EMINIT_PATH, ["-I", f"learnedbest:{NET}", "-r", seed, "-c", K, "-o",
                                         tmp_path / "b.cmp", tmp_path / "x.bin"])
    assert rc == 0 and "judged against k-means++" in out, out
    for a, b in zip(formats.read_cmp(tmp_path / "b.cmp"), refj[:3]):
        np.testing.assert_array_equal(a, b)
    m = re.search(r"final LL (\S+) from the network, (\S+) from k-means\+\+", out)
    assert m, out
    assert [float(m.group(1)), float(m.group(2))] == [refj[4][0], refj[4][1]]
    assert ("kept k-means++" in out) == (refj[3] == "kmeans++")
    assert (refj[3] == "learned") == (refj[4][0] >= refj[4][1])


def test_judge_keeps_higher_final_loglik(net):
    """fem_setnet_judge returns the converged mixture of the better start: its
    E-step log-likelihood is at least the kept final LL's level and the
    reported LLs order the choice."""
    say("learned judge")
    seed = 20_000_503
    X = fem.gengmm(N, D, K, *GEN, seed)[0]
    with fem.Context(N, D, K) as ctx:
        ctx.upload(X)
        w, mu, cov, winner, (l_learned, l_kpp) = net.judge(ctx, n_sub=2048, seed=seed)
        ctx.set_params(w, mu, cov)
        ctx.factor()
        ll = ctx.estep()
    best = max(l_learned, l_kpp)
    assert winner == ("learned" if l_learned >= l_kpp else "kmeans++")
    assert ll >= best - 1e-9 * abs(best)  # one more EM step does not lose likelihood


def test_em2_learned_needs_fewer_iterations(tmp_path):
    """EM from -I learned and from -I kmeanspp on five fresh mixtures of the
    learned initializer's workload (seeds outside its training and evaluation
    ranges), em2's default stopping rule (|dLL/LL| < 1e-5): the learned
    starts need fewer EM sweeps in total."""
    tot = {"learned": 0, "kmeanspp": 0}
    for seed in range(30_100_000, 30_100_005):
        say(f"em2 iterations seed={seed}")
        X = fem.gengmm(N, D, K, *GEN, seed)[0]
        formats.write_bin(tmp_path / "x.bin", X)
        for name, init in (("learned", f"learned:{NET}"), ("kmeanspp", "kmeanspp")):
            rc, out = run_tool(fem.EM2_PATH, ["-c", K, "-r", seed, "-I", init, "-S", tmp_path / "x.bin"])
            m = re.search(r"(\d+) total EM sweeps", out)
            assert rc == 0 and m, out
            tot[name] += int(m.group(1))
    say("total EM sweeps", tot)
    assert tot["learned"] < tot["kmeanspp"]


def test_learned_shape_mismatch(tmp_path):
    X = fem.gengmm(5000, 8, 3, *GEN, 1)[0]
    formats.write_bin(tmp_path / "x.bin", X)
    rc, out = run_tool(fem.EM2_PATH, ["-c", 3, "-I", f"learned:{NET}", "-S", tmp_path / "x.bin"])
    assert "Invalid program argument: learned initializer: the network's d and K differ" in out, out
    rc, out = run_tool(fem.EM2_PATH, ["-c", 3, "-I", "learned:/nonexistent.bin", "-S", tmp_path / "x.bin"])
    assert "Cannot open file with the set network" in out, out
