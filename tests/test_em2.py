"""The em2 driver: argument handling (CPU) and full runs against the oracle (GPU).

The reference's em2 prints LogL with %10.6f; this build also prints
"iter: <n> LogL(17g): <value>" at -v 3 so trajectories can be compared at
1e-10 relative.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import fastem_amd as fem
from fastem_amd import formats

EM2 = fem.EM2_PATH


def run_em2(args, timeout=300):
    p = subprocess.run([EM2] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stdout + p.stderr


@pytest.fixture(scope="module")
def small_bin(tmp_path_factory):
    d = tmp_path_factory.mktemp("em2")
    X = fem.gengmm(3000, 4, 3, -8.0, 8.0, 0.1, 21)[0]
    path = str(d / "small.bin")
    formats.write_bin(path, X)
    return path, X


@pytest.mark.parametrize("args,msg", [
    (["-c", "0"], "Invalid program argument: %invalid value of -c option"),
    (["-c", "2", "-m", "-1"], "Invalid program argument: invalid value of -m option"),
    (["-c", "2", "-a", "0.5"], "argument to -a must be greater then 1.0"),
    (["-c", "2", "-f", "0"], "argument to -f must be greater or than 0"),
    (["-c", "2", "-g", "1.5"], "argument to -g must be small positive number"),
    (["-c", "2", "-A", "bogus"], "Invalid program argument: Unknown EM algorithm"),
    (["-c", "2", "-R", "bogus"], "Invalid program argument: Unknown OpenMP reducer name"),
    (["-c", "2", "-N", "foo"], "Invalid program argument: Bad value of MStepLoopParam"),
    (["-c", "2", "-I", "/nonexistent.cmp"], "Cannot open file with Gaussian mixture"),
])
def test_em2_argument_errors(small_bin, args, msg):
    rc, out = run_em2(args + [small_bin[0]])
    assert rc == 0
    assert msg in out


@pytest.mark.parametrize("n,d,K,spec", [(20000, 16, 50, "automax:1"), (20000, 16, 50, "automax:18"),
                                        (30000, 8, 200, "automax:2"), (5000, 4, 3, "7:3"),
                                        (5000, 4, 3, "2:9"), (5000, 4, 3, None)])
def test_em2_block_counts(orc, tmp_path, n, d, K, spec):
    """em2 -N prints max(nDens, nMStep) blocks (EM2/main.cpp:60-69) as the
    oracle's EM2/BlockManager.cpp:10-28 computes them.  The run stops at the
    missing initializer file, before any device work."""
    X = np.zeros((n, d), dtype=np.float32)
    formats.write_bin(tmp_path / "x.bin", X)
    args = ["-c", K, "-I", "/nonexistent.cmp"] + (["-N", spec] if spec else []) + [tmp_path / "x.bin"]
    rc, out = run_em2(args)
    m = re.search(r"Loading done in (\d+) blocks", out)
    assert m, out
    assert int(m.group(1)) == max(orc.parse_block_params(spec, n, d, K))
    if spec == "automax:1":
        assert int(m.group(1)) > 1


def test_bin_long_header_roundtrip_and_em2_load(tmp_path):
    """The 16-byte .bin header (int32 -1, int32 d, int64 N;
    Util/MultithreadedDataset.cpp:119-134) reads back the same rows, and em2
    loads it as the short form (same row count, features and covariance
    trace; the run stops at the missing initializer file, before device
    work)."""
    X = fem.gengmm(5000, 7, 3, -5.0, 5.0, 0.1, 23)[0]
    formats.write_bin(tmp_path / "s.bin", X)
    formats.write_bin(tmp_path / "l.bin", X, long_header=True)
    raw = open(tmp_path / "l.bin", "rb").read()
    assert raw[:16] == np.array([-1, 7], "<i4").tobytes() + np.array([5000], "<i8").tobytes()
    np.testing.assert_array_equal(formats.read_bin(tmp_path / "l.bin"), X)
    outs = []
    for name in ("s.bin", "l.bin"):
        rc, out = run_em2(["-c", 3, "-I", "/nonexistent.cmp", tmp_path / name])
        assert "5000 vectors, 7 features" in out, out
        outs.append(re.search(r"Covariance matrix trace: (\S+)", out).group(1))
    assert outs[0] == outs[1]


@pytest.mark.gpu
def test_em2_long_header_same_trajectory(orc, tmp_path):
    print("em2 long header", flush=True)
    n, d, K = 20000, 6, 4
    # overlapping components: EM moves for all 8 sweeps (well separated ones
    # reach a fixed point, and em2 stops, after 3)
    X = fem.gengmm(n, d, K, -1.5, 1.5, 0.5, 24)[0]
    w0, mu0, cov0 = orc.kmeanspp_init(X, K, seed=3)
    formats.write_cmp(tmp_path / "init.cmp", w0, mu0, cov0)
    traces = []
    for name, long_header in (("s.bin", False), ("l.bin", True)):
        formats.write_bin(tmp_path / name, X, long_header=long_header)
        rc, out = run_em2(["-c", K, "-I", tmp_path / "init.cmp", "-m", 8, "-f", "1e-300", "-S", "-v", "3",
                           tmp_path / name])
        assert rc == 0, out
        traces.append(parse_trace(out))
    assert len(traces[0]) == 8
    np.testing.assert_array_equal(traces[0], traces[1])


def test_em2_usage_without_file():
    rc, out = run_em2([])
    assert "Arguments: filename" in out


def test_formats_roundtrip(tmp_path):
    X = np.arange(24, dtype=np.float32).reshape(6, 4)
    formats.write_bin(tmp_path / "a.bin", X)
    assert np.array_equal(formats.read_bin(tmp_path / "a.bin"), X)
    rng = np.random.default_rng(0)
    w, mu, cov = rng.random(3), rng.random((3, 2)), rng.random((3, 2, 2))
    formats.write_cmp(tmp_path / "a.cmp", w, mu, cov)
    w2, mu2, cov2 = formats.read_cmp(tmp_path / "a.cmp")
    assert np.array_equal(w, w2) and np.array_equal(mu, mu2) and np.array_equal(cov, cov2)
    raw = open(tmp_path / "a.cmp", "rb").read()
    assert len(raw) == 8 + 3 * (8 * 4 + 8 * 2 + 8) + 8


def parse_trace(out):
    return np.array([float(v) for v in re.findall(r"LogL\(17g\): (\S+)", out)])


@pytest.mark.gpu
@pytest.mark.parametrize("algo,g", [("matrixem", 0.0), ("cachedem", 0.0),
                                    ("simplematrixem", 0.02)])
def test_em2_matches_oracle(orc, tmp_path, algo, g):
    """em2 -I init.cmp on config-1 data (N=100,000, d=16, K=8): per-iteration LL
    and the written mixture agree with the oracle to 1e-10."""
    print(f"em2 {algo} g={g}", flush=True)
    n, d, K, iters = 100000, 16, 8, 25
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 31)[0]
    w0, mu0, cov0 = orc.kmeanspp_init(X, K, seed=4)
    formats.write_bin(tmp_path / "x.bin", X)
    formats.write_cmp(tmp_path / "init.cmp", w0, mu0, cov0)
    rc, out = run_em2(["-c", K, "-A", algo, "-R", "log2", "-I", tmp_path / "init.cmp", "-m", iters,
                       "-f", "1e-300", "-g", g, "-S", "-v", "3", "-o", tmp_path / "out.cmp",
                       tmp_path / "x.bin"])
    assert rc == 0, out
    trace = parse_trace(out)
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=iters, eps=1e-300, g=g, svd=False)
    assert len(trace) == iters
    np.testing.assert_allclose(trace, ref.ll_trace, rtol=1e-10)
    assert f"{iters} total EM sweeps" in out
    w, mu, cov = formats.read_cmp(tmp_path / "out.cmp")
    for a, b in ((w, ref.w), (mu, ref.mu), (cov, ref.cov)):
        assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b))


@pytest.mark.gpu
def test_em2_burnin_and_svd(orc, tmp_path):
    """-b restarts from the same init (the timed run equals a plain run); the
    default condition-number check runs each iteration."""
    print("em2 burn-in", flush=True)
    n, d, K = 20000, 6, 4
    X = fem.gengmm(n, d, K, -10.0, 10.0, 0.1, 32)[0]
    w0, mu0, cov0 = orc.kmeanspp_init(X, K, seed=2)
    formats.write_bin(tmp_path / "x.bin", X)
    formats.write_cmp(tmp_path / "init.cmp", w0, mu0, cov0)
    common = ["-c", K, "-I", tmp_path / "init.cmp", "-m", 10, "-f", "1e-300", "-v", "3"]
    rc, out_b = run_em2(common + ["-b", 3, tmp_path / "x.bin"])
    assert rc == 0, out_b
    assert "Running 3 burn-in iterations" in out_b
    rc, out_p = run_em2(common + [tmp_path / "x.bin"])
    tb, tp = parse_trace(out_b), parse_trace(out_p)
    assert len(tb) == 3 + 10 and len(tp) == 10
    np.testing.assert_array_equal(tb[3:], tp)
    rc, out_v2 = run_em2(["-c", K, "-I", tmp_path / "init.cmp", "-m", 3, "-f", "1e-300", "-v",
                          "2", tmp_path / "x.bin"])
    assert "Cond:" in out_v2 and "Trace:" in out_v2
    ref = orc.train(X, w0, mu0, cov0, maxsweeps=10, eps=1e-300, svd=True)
    np.testing.assert_allclose(tp, ref.ll_trace, rtol=1e-10)


@pytest.mark.gpu
def test_em2_condition_abort(orc, tmp_path):
    print("em2 condition abort", flush=True)
    n, d, K = 5000, 3, 2
    X = fem.gengmm(n, d, K, -5.0, 5.0, 0.1, 33)[0]
    w0, mu0, cov0 = orc.kmeanspp_init(X, K, seed=1)
    cov0[0] = np.diag([1.0, 1.0, 1e-4])
    formats.write_bin(tmp_path / "x.bin", X)
    formats.write_cmp(tmp_path / "init.cmp", w0, mu0, cov0)
    rc, out = run_em2(["-c", K, "-I", tmp_path / "init.cmp", "-a", "1.5", tmp_path / "x.bin"])
    assert "EM algorithm failed: Condition number of a covariance matrix is too high" in out
    bad = cov0.copy()
    bad[1] = np.array([[1.0, 2.0, 0], [2.0, 1.0, 0], [0, 0, 1.0]])
    formats.write_cmp(tmp_path / "bad.cmp", w0, mu0, bad)
    rc, out = run_em2(["-c", K, "-I", tmp_path / "bad.cmp", "-S", tmp_path / "x.bin"])
    # the reference throws while reading a non-PD mixture file (GaussianMixture(FILE*))
    assert rc != 0 or "Cholesky" in out
