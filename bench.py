"""Benchmark: EM iterations per second on one MI355X.

Workload: plain full-covariance GMM EM with FP64 sufficient statistics,
N=4,194,304 points (FP32), d=64, K=64; data from the seeded generator (means
U[-20,20]^d, covariances A A^T/d + 0.05 I, weights Dirichlet(1)), k-means++
initialisation.  One step = one EM iteration on the device: E-step (fused
log-densities + log-sum-exp + posteriors), M-step means, M-step covariances,
regularisation and Cholesky/log-det of the new covariances.

Prints one JSON line.  The CPU baseline is a blocked FP64 restatement of the
reference's SimpleMatrixEM (BLAS on row blocks, one worker per host core)
timed on a quarter of the same workload's rows on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, D, K = 4194304, 64, 64
# Dataset seeds of the benchmark mixture and of the two further time-to-loglik
# datasets.  They lie outside the seed ranges the learned initializer was
# trained on (1..20,000) and evaluated on (10,000,000.., 20,000,000..), so the
# learned runs see unseen mixtures.
SEED = 30_000_001
TTL_EXTRA_SEEDS = (30_000_002, 30_000_003)
PEAK_FP64_MFMA = 78.6e12   # MI355X dense FP64 matrix, FLOP/s
PEAK_FP32_MFMA = 157.3e12  # MI355X dense FP32 matrix, FLOP/s
PEAK_BF16_MFMA = 2.5e15    # MI355X dense BF16 matrix, FLOP/s
PEAK_HBM = 8.0e12          # B/s


def algorithmic_flops(n, d, k):
    """Contractions only (LAWN-41): TRMM N K d^2, GEMM 2 N K d, SYRK N K d(d+1)."""
    return {"estep": n * k * d * d, "means": 2 * n * k * d, "cov": n * k * d * (d + 1)}


def cpu_baseline(X, w0, mu0, cov0, rows, iters):
    """Times the reference's algorithm shape on this host's cores: the blocked
    FP64 restatement of SimpleMatrixEM (dtrmm / dgemm / dsyrk on row blocks
    sized by -N automax:<L3 MB>, one worker per core, oracle/blocked_em.py) on
    the first `rows` rows (all N by default: about 13 s of CPU work on 16
    cores), one burn-in iteration then `iters` timed ones, scaled to the full
    N when rows < N (the work is linear in N)."""
    from oracle import blocked_em as be
    host = be.host_description()
    workers = max(1, min(host["affinity"], 16))
    l3 = host["l3_mb_per_cpu0"] or 32.0
    nblocks = be.automax_blocks(rows, D, K, l3)
    sub = np.ascontiguousarray(X[:rows])
    done, t = 0, 0.0
    with be.BlockedEM(sub, K, threads=workers, nblocks=nblocks) as em:
        try:
            w, mu, cov = em.iterate(w0, mu0, cov0)[1:]  # burn-in
            for _ in range(iters):
                t0 = time.perf_counter()
                w, mu, cov = em.iterate(w, mu, cov)[1:]
                t += time.perf_counter() - t0
                done += 1
        except np.linalg.LinAlgError:  # a covariance of the sample lost definiteness
            pass
    if done == 0:
        return {"value": None, "unit": "EM iters/s", "cores": workers, "kind": "port", "host": host,
                "sample": "no timed iteration: a covariance of the sample was not positive definite"}
    iters = done
    per_iter_full = (t / iters) * (N / rows)
    rows_text = f"all {N} rows" if rows == N else f"the first {rows} of {N} rows, scaled by {N}/{rows}"
    return {"value": 1.0 / per_iter_full, "unit": "EM iters/s", "cores": workers, "kind": "port",
            "host": host,
            "sample": f"{iters} FP64 EM iterations after 1 burn-in of the blocked SimpleMatrixEM "
                      f"restatement (dtrmm/dgemm/dsyrk through numpy's OpenBLAS, {workers} worker "
                      f"processes, -N automax:{l3:g} -> {nblocks} blocks per partition, naive "
                      f"reducer) on {rows_text}; {t:.1f} s measured"}


def em2_modes(X, w0, mu0, cov0, iters=10, log=print):
    """em2 at the benchmark workload in its default mode (condition-number
    check of every covariance after each factorisation, EM2/EMAlgorithm.cpp:
    147-151,159-163) and with -S (check off): em2's own "Iteration time"."""
    import re
    import subprocess
    import tempfile
    import fastem_amd as fem
    from fastem_amd import formats
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        xb, ic = os.path.join(tmp, "x.bin"), os.path.join(tmp, "init.cmp")
        formats.write_bin(xb, X)
        formats.write_cmp(ic, w0, mu0, cov0)
        for name, extra in (("default", []), ("no_svd", ["-S"])):
            log(f"em2 {name}")
            p = subprocess.run([fem.EM2_PATH, "-c", str(K), "-I", ic, "-m", str(iters), "-f", "1e-300",
                                "-v", "0"] + extra + [xb], capture_output=True, text=True, timeout=600)
            m = re.search(r"Iteration time: (\S+) miliseconds", p.stdout)
            out[name + "_ms_per_iter"] = float(m.group(1)) if m else None
            if not m:
                out[name + "_tail"] = (p.stdout + p.stderr)[-400:]
    a, b = out.get("default_ms_per_iter"), out.get("no_svd_ms_per_iter")
    out["check_overhead_frac"] = (a - b) / b if a and b else None
    out["iters"] = iters
    return out


def batch_workload(nfits=256, seed=7, maxsweeps=100, eps=1e-6, g=1e-3, compare=False, log=print,
                   seedings=(1, 1)):
    """Many independent fits advanced together (fem_batch_*): nfits fits in
    groups of 8 sharing a dataset (N log-uniform in [1e4, 1e6], d in
    {8, 16, 32, 64}, true K uniform in [4, 64]); each group is a K sweep over
    four orders around the true K with two k-means++ restarts each (device
    k-means++ + Lloyd with different seeds), scored by BIC; covariance
    regularisation g as the reference's -g.  One persistent launch per EM
    phase over all fits, iterations replayed from a captured hipGraph, launch
    shapes from the autotuner.  Reports fit-iterations/s and algorithmic
    GFLOP/s of the batched EM (the initialisations are not timed)."""
    import fastem_amd as fem
    rng = np.random.default_rng(seed)
    groups = nfits // 8
    fits, data, truth = [], [], []
    t = time.perf_counter()
    for gi in range(groups):
        n = int(round(np.exp(rng.uniform(np.log(1e4), np.log(1e6)))))
        d = int(rng.choice([8, 16, 32, 64]))
        kt = int(rng.integers(4, 65))
        X, w_true, mu_true, cov_true = fem.gengmm(n, d, kt, -10.0, 10.0, 0.1, 1000 + gi)
        data.append(X)
        truth.append((w_true, mu_true, cov_true))
        for K in (max(4, kt - 2), kt, min(64, kt + 2), min(64, kt + 4)):
            with fem.Context(n, d, K) as ctx:
                ctx.upload(X)
                for r in range(2):
                    try:
                        init = ctx.kmeans_init(seed * 1000003 + gi * 16 + K * 2 + r, seeding=seedings[r],
                                               max_sweeps=29)
                    except fem.FemError:
                        init = ctx.get_params()   # mixture written before the error
                    fits.append((gi, n, d, K, kt, init))
    t_init = time.perf_counter() - t
    log(f"batch: {len(fits)} fits, {sum(f[1] for f in fits)} points in total")
    b = fem.Batch([(f[1], f[2], f[3]) for f in fits])
    for i, f in enumerate(fits):
        b.upload(i, data[f[0]])
        b.set_params(i, *f[5])
    t = time.perf_counter()
    report = b.tune()
    t_tune = time.perf_counter() - t
    for i, f in enumerate(fits):
        b.set_params(i, *f[5])
    t = time.perf_counter()
    b.run(maxsweeps, eps=eps, g=g, use_graph=True)
    t_run = time.perf_counter() - t
    res = [b.result(i) for i in range(len(fits))]
    b.close()
    # diagnostic (untimed): each dataset's true-K fit started from the
    # generating mixture, one batch of one fit per dataset
    bt = fem.Batch([(data[gi].shape[0], data[gi].shape[1], truth[gi][0].shape[0]) for gi in range(groups)])
    for gi in range(groups):
        bt.upload(gi, data[gi])
        bt.set_params(gi, *truth[gi])
    bt.run(maxsweeps, eps=eps, g=g, use_graph=True)
    res_true = [bt.result(gi) for gi in range(groups)]
    bt.close()
    iters = np.array([r["iter_counter"] for r in res])
    flop = sum(it * f[1] * f[3] * (2 * f[2] * f[2] + 3 * f[2]) for it, f in zip(iters, fits))
    bic_ok, scored, kdiff, bic_true_ok, kt_better = 0, 0, [], 0, 0
    for gi in range(groups):
        cand = [(fem.bic(r["ll"], f[1], f[2], f[3]), f[3]) for r, f in zip(res, fits)
                if f[0] == gi and r["status"] != fem.Batch.FAILED]
        if cand:
            kt = next(f[4] for f in fits if f[0] == gi)
            scored += 1
            bic_ok += int(min(cand)[1] == kt)
            kdiff.append(min(cand)[1] - kt)
            rt = res_true[gi]
            if rt["status"] != fem.Batch.FAILED:
                bt_true = fem.bic(rt["ll"], data[gi].shape[0], data[gi].shape[1], kt)
                # true K from the generating mixture against the k-means++ fits of the other orders
                others = [c for c in cand if c[1] != kt]
                bic_true_ok += int(not others or bt_true < min(others)[0])
                best_kt = [c[0] for c in cand if c[1] == kt]
                kt_better += int(bool(best_kt) and bt_true < min(best_kt))
    out = {"fits": len(fits), "datasets": groups, "points_total": int(sum(f[1] for f in fits)),
           "seconds": t_run, "tune_seconds": t_tune, "init_seconds_untimed": t_init,
           "fit_iterations": int(iters.sum()),
           "fit_iterations_per_s": float(iters.sum() / t_run), "gflops_algorithmic": flop / t_run / 1e9,
           "gflops_note": "dense-equivalent: the contraction flops of every (point, component) pair, "
                          "while the covariance pass sums over the nonzero posteriors only",
           "converged": int(sum(r["status"] == fem.Batch.CONVERGED for r in res)),
           "failed": int(sum(r["status"] == fem.Batch.FAILED for r in res)),
           "bic_true_k_recovered": f"{bic_ok}/{scored}",
           "bic_true_k_recovered_true_k_fit_from_generating_mixture": f"{bic_true_ok}/{scored}",
           "true_k_fit_from_generating_mixture_beats_kmeanspp_fits": f"{kt_better}/{scored}",
           "bic_k_minus_true_k": {str(v): int(sum(1 for x in kdiff if x == v)) for v in sorted(set(kdiff))},
           "eps": eps, "g": g,
           "autotuner": report.strip().splitlines()[-1] if report else None}
    if compare:
        # the same fits one at a time on single-fit contexts, same iteration counts
        t = time.perf_counter()
        for f, it in zip(fits, iters):
            with fem.Context(f[1], f[2], f[3]) as ctx:
                ctx.upload(data[f[0]])
                ctx.set_regularization(g)
                ctx.set_params(*f[5])
                try:
                    ctx.factor()
                    for _ in range(int(it)):
                        ctx.estep()
                        ctx.mstep(copy=False)
                        ctx.factor()
                except fem.FemError:
                    pass
        out["sequential_single_fit_seconds"] = time.perf_counter() - t
        out["sequential_note"] = ("same fits and iteration counts on single-fit contexts, "
                                  "including per-fit context creation and upload")
    return out


def precision_variant(iters=8, seed=2025, log=print):
    """The precision workload: N = 4,194,304 FP32 points, d = 128, K = 128, the
    benchmark's generator (means U[-20, 20], covariances A A^T/d + 0.05 I).
    One device k-means++ init, then `iters` EM iterations in each form from
    it on the same context; per iteration the FP32 and BF16x3 forms' LL
    (E-step) and weights, means and covariances (M-step) against the FP64
    form's, as max-norm relative errors, and each form's GPU time per
    iteration (HIP events of the four phases).  The FP64 form matches the
    FP64 oracle to 1e-10 in the parity tests, so it stands in for the oracle
    at this size."""
    import fastem_amd as fem
    n, d, k = 4194304, 128, 128
    X = fem.gengmm(n, d, k, -20.0, 20.0, 0.05, seed)[0]
    out = {"n": n, "d": d, "k": k, "iters": iters, "reference": "fp64 form"}
    with fem.Context(n, d, k, fem.PREC_FP64, 0) as ctx:
        ctx.upload(X)
        del X
        try:
            init = ctx.kmeans_init(seed=1, seeding=1, max_sweeps=29)
        except fem.FemError:
            init = ctx.get_params()   # the mixture written before the error
        traj = {}
        for name, prec in (("fp64", fem.PREC_FP64), ("fp32", fem.PREC_FP32), ("bf16x3", fem.PREC_BF16X3)):
            log(f"precision workload: {name}")
            ctx.set_precision(prec)
            ctx.set_params(*init)
            ctx.factor()
            lls, params, ms, ems = [], [], [], []
            for _ in range(iters):
                lls.append(ctx.estep())
                t = ctx.times()
                params.append(ctx.mstep(copy=True))
                t2 = ctx.times()
                ctx.factor()
                t3 = ctx.times()
                ms.append(t["estep_ms"] + t2["means_ms"] + t2["cov_ms"] + t3["factor_ms"])
                ems.append(t["estep_ms"])
            traj[name] = (lls, params)
            out[f"{name}_gpu_ms_per_iter"] = float(np.median(ms))
            out[f"{name}_estep_tflops"] = algorithmic_flops(n, d, k)["estep"] / (np.median(ems) * 1e-3) / 1e12
        ref_ll, ref_p = traj["fp64"]

        def rel(a, b):
            return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))

        def rel_per_component(a, b):
            # worst component of max|a_k - b_k| / max|b_k|
            a, b = np.asarray(a), np.asarray(b)
            axes = tuple(range(1, a.ndim))
            return float(np.max(np.max(np.abs(a - b), axis=axes) / np.max(np.abs(b), axis=axes)))

        for name in ("fp32", "bf16x3"):
            lls, params = traj[name]
            out[f"{name}_rel_err_per_iter"] = {
                "ll": [abs(a - b) / abs(b) for a, b in zip(lls, ref_ll)],
                "w": [rel(p[0], q[0]) for p, q in zip(params, ref_p)],
                "mu": [rel(p[1], q[1]) for p, q in zip(params, ref_p)],
                "cov": [rel(p[2], q[2]) for p, q in zip(params, ref_p)],
                "mu_worst_component": [rel_per_component(p[1], q[1]) for p, q in zip(params, ref_p)],
                "cov_worst_component": [rel_per_component(p[2], q[2]) for p, q in zip(params, ref_p)]}
        out["fp64_cov_max_abs_per_iter"] = [float(np.max(np.abs(q[2]))) for q in ref_p]
        out["fp64_cov_change_per_iter"] = [rel(b[2], a[2]) for a, b in zip(ref_p[:-1], ref_p[1:])]
        out["fp64_ll_per_iter"] = ref_ll
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a 1 GiB copy); None
    if the summary does not list it."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_n{N}_d{D}_k{K}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.rstrip("\n").split("\t")
        if not line.startswith("#") and len(parts) == 4 and parts[0] == kernel:
            return float(parts[3])
    return None


def calib_copy():
    """A plain device-to-device copy of exactly 1 GiB, three times: the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": x.numel() * 4, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP32-form measurement")
    ap.add_argument("--cpu-rows", type=int, default=N)
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-em2", action="store_true",
                    help="skip the em2 default-mode (condition-number check) measurement")
    ap.add_argument("--no-ttl", action="store_true",
                    help="skip the time-to-log-likelihood measurement")
    ap.add_argument("--no-ttl-extra", action="store_true",
                    help="time-to-loglik on the benchmark dataset only")
    ap.add_argument("--ttl-cap", type=int, default=400,
                    help="sweep cap of the converged FP64 reference run")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the many-fit batched workload")
    ap.add_argument("--batch-fits", type=int, default=256)
    ap.add_argument("--batch-compare", action="store_true",
                    help="also time the batched fits one at a time")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the d=128, K=128 precision workload")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        calib_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine runs a fit on a single MI355X "
              "and has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import fastem_amd as fem

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    log("generating data")
    X, _, _, _ = fem.gengmm(N, D, K, -20.0, 20.0, 0.05, SEED)
    log("k-means++ init")
    t_init = time.perf_counter()
    w0, mu0, cov0 = fem.kmeanspp_init(X, K, seed=1, lloyd_iters=29)
    t_init = time.perf_counter() - t_init

    def run_form(prec, steps, warmup, nonzero=False):
        ctx = fem.Context(N, D, K, prec, 0)
        ctx.upload(X)
        ctx.set_params(w0, mu0, cov0)
        ctx.factor()

        def step():
            ll = ctx.estep()
            t = ctx.times()
            ctx.mstep(copy=False)
            t2 = ctx.times()
            ctx.factor()
            t3 = ctx.times()
            return ll, t["estep_ms"], t2["means_ms"], t2["cov_ms"], t3["factor_ms"]

        for _ in range(warmup):
            step()
        recs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            recs.append(step())
        wall = time.perf_counter() - t0
        # fraction of the last E-step's posteriors that are nonzero: the
        # (point, component) pairs the covariance pass sums over
        nz = float(np.count_nonzero(ctx.posteriors())) / (N * K) if nonzero else None
        skipped = ctx.estep_skipped()  # of the last E-step (0 unless pruned)
        ctx.close()
        ph = {name: float(np.mean([r[i] for r in recs]))
              for i, name in ((1, "estep"), (2, "means"), (3, "cov"), (4, "factor"))}
        return {"iters_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
                "phases_ms": ph, "ll": [recs[0][0], recs[-1][0]], "nonzero": nz, "skipped": skipped}

    def with_env(env, fn):
        old = {k2: os.environ.get(k2) for k2 in env}
        os.environ.update(env)
        try:
            return fn()
        finally:
            for k2, v in old.items():
                if v is None:
                    os.environ.pop(k2, None)
                else:
                    os.environ[k2] = v

    log("fp64 form")
    r64 = run_form(fem.PREC_FP64, args.steps, args.warmup, nonzero=True)
    # the FP64 E-step evaluating every component for every point (no pruning,
    # point order): the dense kernel's roofline and the iteration without the
    # skip; and the iteration with both the E-step and the covariances dense,
    # the reference's shape of work
    log("fp64 form, dense E-step")
    re_d = with_env({"FASTEM_ESTEP_DENSE": "1"}, lambda: run_form(fem.PREC_FP64, args.steps, args.warmup))
    variants = {"fp64_dense_estep": {"iters_per_s": re_d["iters_per_s"], "ms_per_step": re_d["ms_per_step"],
                                     "phases_ms": re_d["phases_ms"], "loglik_first_last": re_d["ll"]}}
    if not args.no_variants:
        log("fp64 form, dense E-step and covariance sum")
        rd = with_env({"FASTEM_ESTEP_DENSE": "1", "FASTEM_COV_DENSE": "1"},
                      lambda: run_form(fem.PREC_FP64, args.steps, args.warmup))
        variants["fp64_dense"] = {"iters_per_s": rd["iters_per_s"], "ms_per_step": rd["ms_per_step"],
                                  "phases_ms": rd["phases_ms"], "loglik_first_last": rd["ll"]}
        log("fp32 form")
        r32 = run_form(fem.PREC_FP32, args.steps, args.warmup)
        variants["fp32"] = {
            "iters_per_s": r32["iters_per_s"], "phases_ms": r32["phases_ms"],
            "estep_frac_of_fp32_mfma_peak":
                algorithmic_flops(N, D, K)["estep"] / (r32["phases_ms"]["estep"] * 1e-3)
                / PEAK_FP32_MFMA}
        log("bf16x3 form")
        rb = run_form(fem.PREC_BF16X3, args.steps, args.warmup)
        # the same form with every (point, component) pair centred on its own
        # mean in the E-step (the reference-centred kernel switched off)
        log("bf16x3 form, component-centred E-step")
        os.environ["FASTEM_ESTEPB_REF"] = "0"
        try:
            rbc = run_form(fem.PREC_BF16X3, args.steps, args.warmup)
        finally:
            del os.environ["FASTEM_ESTEPB_REF"]
        ebf = algorithmic_flops(N, D, K)["estep"]
        variants["bf16x3"] = {
            "iters_per_s": rb["iters_per_s"], "phases_ms": rb["phases_ms"],
            "estep_fp32_equivalent_frac_of_fp32_mfma_peak":
                ebf / (rb["phases_ms"]["estep"] * 1e-3) / PEAK_FP32_MFMA,
            "estep_frac_of_bf16_peak_3_products":
                3 * ebf / (rb["phases_ms"]["estep"] * 1e-3) / PEAK_BF16_MFMA,
            "estep_ms_component_centred": rbc["phases_ms"]["estep"],
            "iters_per_s_component_centred": rbc["iters_per_s"]}

    fl = algorithmic_flops(N, D, K)
    ph = r64["phases_ms"]
    # the pruned E-step's MFMA work per launch: every (wave, component) forms
    # its first 16-row block (10 4x4x4 quads per 16-point group), and the
    # evaluations not skipped the rest (1.0625 N K d^2 in all at d = 64)
    waves, groups = N // 32, 2
    first_block = 10 * groups * 512.0
    full = 1.0625 * 32 * D * D
    executed = waves * K * (first_block + (1.0 - r64["skipped"]) * (full - first_block))
    achieved = executed / (ph["estep"] * 1e-3)
    achieved_dense = fl["estep"] / (re_d["phases_ms"]["estep"] * 1e-3)
    out = {
        "metric": "EM iters/s (plain FP64-statistics GMM EM, N=4194304 d=64 K=64)",
        "value": r64["iters_per_s"],
        "unit": "EM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded gengmm mixture, k-means++ init)",
        "config": {"workload": "gmm-em-full-cov", "n": N, "d": D, "k": K,
                   "precision": "fp64", "init": "kmeans++ + 29 Lloyd sweeps (device)"},
        "roofline": {"bound": "mfma", "kernel": "k_estep64_pruned<4,2>", "achieved": achieved / 1e12,
                     "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_MFMA,
                     "flops_per_launch": executed,
                     "flops_note": "MFMA flops the pruned E-step executes per launch: the first 16-row "
                                   "block of every (32-point wave, component) pair and the whole "
                                   "triangle for the pairs not skipped (estep_skipped_fraction)",
                     "traffic": measured_traffic("fem::k_estep64_pruned<4, 2>"),
                     "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                     "(profiles/hbm_traffic_n4194304_d64_k64.txt); algorithmic "
                                     "bytes: X read 1.07e9 + FP64 posteriors written 2.15e9"},
        "roofline_dense_estep": {"bound": "mfma", "kernel": "k_estep64_w3<4,2>",
                                 "achieved": achieved_dense / 1e12, "peak": PEAK_FP64_MFMA / 1e12,
                                 "unit": "TFLOP/s", "frac": achieved_dense / PEAK_FP64_MFMA,
                                 "flops_per_launch": fl["estep"],
                                 "traffic": measured_traffic("fem::k_estep64_w3<4, 2>"),
                                 "note": "the same E-step evaluating every component for every point "
                                         "(variants.fp64_dense_estep, FASTEM_ESTEP_DENSE=1): N K d^2 flops"},
        "phases_ms": ph,
        "phase_tflops": {"estep_executed": achieved / 1e12, "means": fl["means"] / ph["means"] / 1e9,
                         "cov_over_nonzero_posteriors": fl["cov"] * r64["nonzero"] / ph["cov"] / 1e9},
        "nonzero_posterior_fraction": r64["nonzero"],
        "estep_skipped_fraction": r64["skipped"],
        "work_note": "exact skips of work that adds exactly zero: the E-step skips a component for a "
                     "wave of 32 points when the first 16 rows of U_k^T (x - mu_k) put every point's "
                     "log-density more than 760 below its running maximum (posterior exactly 0; "
                     "outputs bit for bit those of the dense E-step), and the covariance pass sums "
                     "over the nonzero posteriors only; variants.fp64_dense_estep and fp64_dense "
                     "are the same iteration without the first and without both",
        "loglik_first_last": r64["ll"],
        "variants": variants,
        "init_seconds": t_init,
    }
    if not args.no_ttl:
        from fastem_amd import ttl
        log("time-to-loglik: plain FP64 reference run")
        with fem.Context(N, D, K, fem.PREC_FP64, 0) as ctx:
            ctx.upload(X)
            plain = ttl.run_plain(ctx, seed=1, tol=1e-6, cap=args.ttl_cap, conv=1e-12)
            ll_star = plain["ll_final"]
            runs = {}
            chk = ttl.check_every(10)
            variants = [("fp32", [(fem.PREC_FP32, chk)], None),
                        ("bf16x3", [(fem.PREC_BF16X3, chk)], None),
                        ("bf16x3/switch1e-9", [(fem.PREC_BF16X3, 1e-9)], None)]
            net_path = os.path.join(ROOT, "fastem_amd", "data", f"setnet_d{D}_k{K}.bin")
            learned = None
            if os.path.exists(net_path):
                from fastem_amd.setnet import SetNet
                net = SetNet.load(net_path)
                used = []

                def learned(c):
                    used.append(net.init_context(c, X, n_sub=2048, seed=1)[3])

                variants.append(("learned+bf16x3", [(fem.PREC_BF16X3, chk)], learned))
                variants.append(("learned+bf16x3/switch1e-9", [(fem.PREC_BF16X3, 1e-9)], learned))
            for name, stages, init in variants:
                log(f"time-to-loglik: {name} -> FP64 run")
                try:
                    r = ttl.run_schedule(ctx, seed=1, ll_star=ll_star, stages=stages, tol=1e-6,
                                         cap=args.ttl_cap, init=init)
                except fem.FemError as e:   # e.g. a covariance that lost definiteness
                    runs[name] = {"error": str(e)}
                    continue
                runs[name] = {"seconds": r["t_reach"], "init_s": r["t_init"],
                              "reduced_iters": r["stage_iters"][0], "fp64_iters": r["fp64_iters"],
                              "fp64_checks": r.get("fp64_checks", 0), "ll_final": r["trace"][-1][1],
                              "speedup": (plain["t_reach"] / r["t_reach"]
                                          if plain["t_reach"] and r["t_reach"] else None)}
            own = None
            if learned is not None:
                # The learned start may converge to another optimum than the
                # k-means++ one: targets taken from its own converged FP64 run.
                log("time-to-loglik: learned init, plain FP64 reference run")
                lp = ttl.run_plain(ctx, seed=1, tol=1e-6, cap=args.ttl_cap, conv=1e-12, init=learned)
                own = {"ll_star": lp["ll_final"], "plain_fp64_s": lp["t_reach"],
                       "plain_iters": lp["iters_to_reach"], "iters_run": lp["iters"]}
                for name, stages in (("learned+bf16x3", [(fem.PREC_BF16X3, chk)]),
                                     ("learned+bf16x3/switch1e-9", [(fem.PREC_BF16X3, 1e-9)])):
                    log(f"time-to-loglik: {name} -> FP64 run, own target")
                    try:
                        r = ttl.run_schedule(ctx, seed=1, ll_star=lp["ll_final"], stages=stages,
                                             tol=1e-6, cap=args.ttl_cap, init=learned)
                    except fem.FemError as e:
                        own[name] = {"error": str(e)}
                        continue
                    t = r["t_reach"]
                    own[name] = {"seconds": t, "reduced_iters": r["stage_iters"][0],
                                 "fp64_iters": r["fp64_iters"], "fp64_checks": r.get("fp64_checks", 0),
                                 "ll_final": r["trace"][-1][1],
                                 "speedup_vs_plain_kpp": plain["t_reach"] / t if t and plain["t_reach"] else None,
                                 "speedup_vs_plain_learned": lp["t_reach"] / t if t and lp["t_reach"] else None}
            extra = []
            if learned is not None and not args.no_ttl_extra:
                for s in TTL_EXTRA_SEEDS:
                    log(f"time-to-loglik: dataset seed {s}")
                    Xs = fem.gengmm(N, D, K, -20.0, 20.0, 0.05, s)[0]
                    ctx.upload(Xs)

                    def learned_s(c, Xs=Xs):
                        net.init_context(c, Xs, n_sub=2048, seed=1)

                    try:
                        pk = ttl.run_plain(ctx, seed=1, tol=1e-6, cap=args.ttl_cap, conv=1e-12)
                        pl = ttl.run_plain(ctx, seed=1, tol=1e-6, cap=args.ttl_cap, conv=1e-12,
                                           init=learned_s)
                        stages = [(fem.PREC_BF16X3, chk)]
                        r1 = ttl.run_schedule(ctx, seed=1, ll_star=pk["ll_final"], stages=stages,
                                              tol=1e-6, cap=args.ttl_cap, init=learned_s)
                        r2 = ttl.run_schedule(ctx, seed=1, ll_star=pl["ll_final"], stages=stages,
                                              tol=1e-6, cap=args.ttl_cap, init=learned_s)
                    except fem.FemError as e:
                        extra.append({"seed": s, "error": str(e)})
                        continue
                    t1, t2 = r1["t_reach"], r2["t_reach"]
                    extra.append({
                        "seed": s, "ll_star_kpp": pk["ll_final"], "plain_kpp_s": pk["t_reach"],
                        "plain_kpp_iters": pk["iters_to_reach"], "ll_star_learned": pl["ll_final"],
                        "plain_learned_s": pl["t_reach"], "plain_learned_iters": pl["iters_to_reach"],
                        "learned_bf16x3_to_kpp_target_s": t1, "learned_bf16x3_to_own_target_s": t2,
                        "speedup_to_kpp_target": pk["t_reach"] / t1 if t1 and pk["t_reach"] else None,
                        "speedup_own_target_vs_plain_kpp":
                            pk["t_reach"] / t2 if t2 and pk["t_reach"] else None,
                        "speedup_own_target_vs_plain_learned":
                            pl["t_reach"] / t2 if t2 and pl["t_reach"] else None})
                    del Xs
        out["time_to_loglik"] = {
            "tol": 1e-6, "ll_star": ll_star,
            "ll_star_def": "converged FP64 EM from the same k-means++ init "
                           f"(|dLL/LL| < 1e-12 or {args.ttl_cap} sweeps; ran {plain['iters']})",
            "plain_fp64_s": plain["t_reach"], "plain_iters": plain["iters_to_reach"],
            "init_s": plain["t_init"],
            "mixed": runs,
            "schedule": "reduced-precision EM throughout, one FP64 E-step on the current "
                        "parameters every 10 iterations checking the target (its time counted); "
                        "the /switch1e-9 runs instead switch to FP64 iterations once |dLL/LL| < "
                        "1e-9 or the LL decreases; done at the first FP64 E-step within tol of "
                        "ll_star",
            "init": "kmeans++ + up to 29 Lloyd sweeps on the device (learned: the set network's "
                    "proposal on a 2048-point subsample seeds the same Lloyd sweeps), included in "
                    "every time",
            "learned_own_target": own,
            "more_datasets": extra,
            "more_datasets_def": "two further datasets of the same generator (seeds in 'seed'): "
                                 "plain FP64 EM from k-means++ and from the learned init, each run to "
                                 "convergence for its target, and learned init + BF16x3 -> FP64 "
                                 "timed to within tol of either target",
            "learned_own_target_def": "ll_star of converged FP64 EM from the learned init itself "
                                      "(|dLL/LL| < 1e-12 or the sweep cap); the learned runs timed "
                                      "to within tol of it"}
    if not args.no_precision:
        log("precision workload, N=4194304 d=128 K=128")
        try:
            out["precision_d128_k128"] = precision_variant(log=log)
        except fem.FemError as e:
            out["precision_d128_k128"] = {"error": str(e)}
    if not args.no_batch:
        log("batched fits")
        out["batch"] = batch_workload(args.batch_fits, compare=args.batch_compare, log=log)
    if not args.no_em2:
        out["em2_modes"] = em2_modes(X, w0, mu0, cov0, log=log)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(X, w0, mu0, cov0, args.cpu_rows, args.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
