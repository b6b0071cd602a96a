"""Training and evaluation pipeline of the learned initializer.

  python -m fastem_amd.setnet_train --d 32 --K 16 --n 65536 --fits 20000 \\
      --eval 1000 --out fastem_amd/data/setnet_d32_k16.bin --report report.json

1. Targets: `--fits` synthetic mixtures (fem_gengmm, seeds 1..fits) are fitted
   to convergence by the batched device EM engine (fem_batch_*), started from
   the generating parameters; each fit keeps a `--keep`-point subsample of its
   data (normalised) and the converged mixture in the same units.
2. Training: Adam on the set network's HIP forward/backward kernels
   (fem_setnet_step), random signed coordinate permutations as augmentation.
3. Evaluation on `--eval` fresh seeds: EM from the learned initialisation
   (network means -> Lloyd -> cluster mixture) and EM from device k-means++
   (seeding -> Lloyd -> cluster mixture) run in one batch to the reference's
   stopping rule; reported per seed are the iterations to converge, the final
   log-likelihood of both, and the time of each initializer on resident data.
"""
import argparse
import json
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import fastem_amd as fem
from fastem_amd.setnet import SetNet, fit_to_targets, normalise, point_targets, subsample

GEN = {(32, 16): (-10.0, 10.0, 0.1), (64, 64): (-20.0, 20.0, 0.05)}


def _gen(args):
    n, d, K, lo, hi, diag, seed = args
    return fem.gengmm(n, d, K, lo, hi, diag, seed)


def make_targets(seeds, n, d, K, gen, keep, eps=1e-8, maxsweeps=200, chunk=256, g=0.0, log=print):
    """Converged engine fits of the mixtures `seeds` (see module docstring)."""
    lo, hi, diag = gen
    F = len(seeds)
    sub = np.empty((F, keep, d), dtype=np.float32)
    M = np.empty((F, K, d), dtype=np.float32)
    V = np.empty((F, K, d), dtype=np.float32)
    P = np.empty((F, K), dtype=np.float32)
    TI = np.zeros((F, keep), dtype=np.uint8)
    ok = np.zeros(F, dtype=bool)
    iters = np.zeros(F, dtype=np.int32)
    t0 = time.perf_counter()
    batch = None
    with ThreadPoolExecutor(2) as pool:
        for c0 in range(0, F, chunk):
            cs = seeds[c0:c0 + chunk]
            if batch is None or len(cs) != len(batch.shapes):
                if batch is not None:
                    batch.close()
                batch = fem.Batch([(n, d, K)] * len(cs))
            subs = []
            for i, (X, w, mu, cov) in enumerate(pool.map(_gen, [(n, d, K, lo, hi, diag, int(s)) for s in cs])):
                batch.upload(i, X)
                batch.set_params(i, w, mu, cov)
                subs.append(subsample(X, keep, int(cs[i])))
            batch.run(maxsweeps, eps=eps, g=g)
            for i in range(len(cs)):
                r = batch.result(i)
                f = c0 + i
                Z, c, s = normalise(subs[i])
                sub[f] = Z
                iters[f] = r["iter_counter"]
                if r["status"] == fem.Batch.FAILED:
                    continue
                m, v, p = fit_to_targets(r["w"], r["mu"], r["cov"], c, s)
                M[f], V[f], P[f] = m, v, p
                TI[f] = point_targets(Z, m)
                ok[f] = True
            log(f"targets {min(c0 + chunk, F)}/{F}  {time.perf_counter() - t0:.1f} s")
    if batch is not None:
        batch.close()
    return {"sub": sub[ok], "M": M[ok], "V": V[ok], "P": P[ok], "TI": TI[ok], "iters": iters[ok],
            "failed": int((~ok).sum()), "seconds": time.perf_counter() - t0}


def augment(rng, Xs, M, V):
    """Random signed permutation of the coordinates, one per set."""
    B, _, d = Xs.shape
    Xa, Ma, Va = np.empty_like(Xs), np.empty_like(M), np.empty_like(V)
    for b in range(B):
        perm = rng.permutation(d)
        sign = np.where(rng.random(d) < 0.5, -1.0, 1.0).astype(np.float32)
        Xa[b] = Xs[b][:, perm] * sign
        Ma[b] = M[b][:, perm] * sign
        Va[b] = V[b][:, perm]
    return Xa, Ma, Va


def gather_targets(M, TI):
    """Point targets T[b, i] = M[b, TI[b, i]]."""
    return np.take_along_axis(M, TI[:, :, None].astype(np.int64), axis=1)


def train(net, data, steps, batch=64, lr=2e-3, gam=1.0, seed=0, val_frac=0.02, log=print, log_every=200):
    rng = np.random.default_rng(seed)
    F = data["sub"].shape[0]
    nval = max(1, int(F * val_frac))
    tr = np.arange(nval, F)
    vb = np.arange(min(nval, 256))

    def val_loss():
        return net.step(data["sub"][vb], data["M"][vb], data["V"][vb], data["P"][vb],
                        T=gather_targets(data["M"][vb], data["TI"][vb]), gam=gam) / len(vb)

    hist = [(0, val_loss())]
    t0 = time.perf_counter()
    for it in range(1, steps + 1):
        idx = rng.choice(tr, size=batch, replace=False)
        Xs, M, V = augment(rng, data["sub"][idx], data["M"][idx], data["V"][idx])
        lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * (it - 1) / steps))
        net.step(Xs, M, V, data["P"][idx], T=gather_targets(M, data["TI"][idx]), gam=gam, lr=lr_t)
        if it % log_every == 0 or it == steps:
            hist.append((it, val_loss()))
            log(f"step {it}/{steps} val loss {hist[-1][1]:.4f}  {time.perf_counter() - t0:.1f} s")
    return {"val_loss": hist, "seconds": time.perf_counter() - t0}


def evaluate(net, seeds, n, d, K, gen, n_sub=1024, eps=1e-5, maxsweeps=1000, lloyd_iters=29, chunk=128,
             g=0.0, log=print, bf16=True):
    """EM from the learned initialisation vs EM from k-means++ on fresh seeds.

    Both start with Lloyd sweeps and the cluster mixture on the device; they
    differ in the seeds: the network's means (fem_kmeans_init_seeded) or the
    k-means++ seeding pass (fem_kmeans_init).  EM then runs to the reference's
    stopping rule |dLL/LL| < eps for both in one batch.  The learned start is
    the library's proposal (BF16 inference, fem_setnet_propose); bf16 selects
    the forward of the seed-query head's proposal (False: the FP32 forward)."""
    lo, hi, diag = gen
    recs = []
    ctx = fem.Context(n, d, K)
    ctx.set_regularization(g)  # the proposal then keeps rank-deficient clusters when g > 0
    for c0 in range(0, len(seeds), chunk):
        cs = seeds[c0:c0 + chunk]
        inits = []
        for s in cs:
            X, _, _, _ = fem.gengmm(n, d, K, lo, hi, diag, int(s))
            ctx.upload(X)
            t = time.perf_counter()
            try:
                kpp = ctx.kmeans_init(int(s), seeding=1, max_sweeps=lloyd_iters)
            except fem.FemError:
                kpp = None
            t_kpp = time.perf_counter() - t
            t = time.perf_counter()
            try:
                w, mu, cov, used = net.init_context(ctx, n_sub=n_sub, seed=int(s), lloyd_sweeps=lloyd_iters)
                lrn = (w, mu, cov)
            except fem.FemError:
                lrn, used = None, "failed"
            t_lrn = time.perf_counter() - t
            try:
                qry = ctx.kmeans_init_seeded(net.propose_query_means(X, n_sub=n_sub, seed=int(s), bf16=bf16),
                                             seed=int(s), max_sweeps=lloyd_iters)
            except fem.FemError:
                qry = None
            inits.append((X, kpp, lrn, used, t_kpp, t_lrn, qry))
        b = fem.Batch([(n, d, K)] * (3 * len(cs)))
        for i, (X, kpp, lrn, _, _, _, qry) in enumerate(inits):
            cands = [kpp, lrn, qry]
            fallback = next((c for c in cands if c is not None), None)
            if fallback is None:
                continue
            for j, c in enumerate(cands):
                b.upload(3 * i + j, X)
                b.set_params(3 * i + j, *(c if c is not None else fallback))
        b.run(maxsweeps, eps=eps, g=g)
        for i, s in enumerate(cs):
            rk, rl, rq = b.result(3 * i), b.result(3 * i + 1), b.result(3 * i + 2)
            kpp_ok = inits[i][1] is not None and rk["status"] != fem.Batch.FAILED
            lrn_ok = inits[i][2] is not None and rl["status"] != fem.Batch.FAILED
            qry_ok = inits[i][6] is not None and rq["status"] != fem.Batch.FAILED
            recs.append({"seed": int(s), "kpp_ok": bool(kpp_ok), "kpp_ll": rk["ll"] if kpp_ok else None,
                         "kpp_iters": rk["iter_counter"] if kpp_ok else None,
                         "learned_ok": bool(lrn_ok), "learned_ll": rl["ll"] if lrn_ok else None,
                         "learned_iters": rl["iter_counter"] if lrn_ok else None,
                         "learned_used": inits[i][3],
                         "query_ok": bool(qry_ok), "query_ll": rq["ll"] if qry_ok else None,
                         "query_iters": rq["iter_counter"] if qry_ok else None,
                         "kpp_init_s": inits[i][4], "learned_init_s": inits[i][5]})
        b.close()
        log(f"eval {min(c0 + chunk, len(seeds))}/{len(seeds)}")
    ctx.close()
    return recs


def summarise(recs, tol=1e-6):
    both = [r for r in recs if r["kpp_ok"] and r["learned_ok"]]
    rel = np.array([(r["learned_ll"] - r["kpp_ll"]) / abs(r["kpp_ll"]) for r in both])
    best = [max(r["learned_ll"] if r["learned_ok"] else -np.inf, r["kpp_ll"] if r["kpp_ok"] else -np.inf)
            for r in recs]
    kpp_ll = [r["kpp_ll"] if r["kpp_ok"] else -np.inf for r in recs]
    return {
        "seeds": len(recs),
        "kpp_failed": sum(not r["kpp_ok"] for r in recs),
        "learned_failed": sum(not r["learned_ok"] for r in recs),
        "mean_iters_kpp": float(np.mean([r["kpp_iters"] for r in both])) if both else None,
        "mean_iters_learned": float(np.mean([r["learned_iters"] for r in both])) if both else None,
        "median_iters_kpp": float(np.median([r["kpp_iters"] for r in both])) if both else None,
        "median_iters_learned": float(np.median([r["learned_iters"] for r in both])) if both else None,
        "learned_better": int((rel > tol).sum()),
        "same_optimum": int((np.abs(rel) <= tol).sum()),
        "learned_worse": int((rel < -tol).sum()),
        "mean_rel_ll_learned_minus_kpp": float(rel.mean()) if len(rel) else None,
        "best_of_improves_on_kpp": int(sum(b > k + tol * abs(k) for b, k in zip(best, kpp_ll)
                                           if np.isfinite(k))),
        "mean_init_s_kpp": float(np.mean([r["kpp_init_s"] for r in recs])),
        "mean_init_s_learned": float(np.mean([r["learned_init_s"] for r in recs])),
        "learned_fell_back_to_kpp": sum(r["learned_used"] != "learned" for r in recs),
        "query_head": _vs_kpp(recs, "query", tol),
    }


def _vs_kpp(recs, key, tol):
    both = [r for r in recs if r["kpp_ok"] and r[key + "_ok"]]
    rel = np.array([(r[key + "_ll"] - r["kpp_ll"]) / abs(r["kpp_ll"]) for r in both])
    return {"better": int((rel > tol).sum()), "same": int((np.abs(rel) <= tol).sum()),
            "worse": int((rel < -tol).sum()),
            "mean_iters": float(np.mean([r[key + "_iters"] for r in both])) if both else None}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--d", type=int, default=32)
    ap.add_argument("--K", type=int, default=16)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--fits", type=int, default=20000)
    ap.add_argument("--keep", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--gam", type=float, default=1.0, help="weight of the point-head loss")
    ap.add_argument("--eval", type=int, default=1000)
    ap.add_argument("--eval-seed0", type=int, default=10_000_000)
    ap.add_argument("--eval-eps", type=float, default=1e-5)
    ap.add_argument("--eval-n", type=int, default=None, help="points per evaluation dataset (default --n)")
    ap.add_argument("--eval-chunk", type=int, default=128,
                    help="evaluation seeds per batched EM run (3 fits each; bounds device memory)")
    ap.add_argument("--g", type=float, default=0.0,
                    help="covariance regularisation of the target fits and of the evaluation EM")
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--net-seed", type=int, default=0)
    ap.add_argument("--load", default=None,
                    help="evaluate a saved network (.bin) instead of generating targets and training")
    ap.add_argument("--fp32-inference", action="store_true",
                    help="evaluate with the FP32 forward instead of the BF16-MFMA inference forward")
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, flush=True)

    gen = GEN.get((a.d, a.K), (-10.0, 10.0, 0.1))
    if a.load:
        net = SetNet.load(a.load)
        assert (net.d, net.K) == (a.d, a.K), "saved network shape differs from --d/--K"
        data, tr = None, None
    else:
        log(f"targets: {a.fits} fits n={a.n} d={a.d} K={a.K} gen={gen}")
        data = make_targets(np.arange(1, a.fits + 1), a.n, a.d, a.K, gen, a.keep, g=a.g, log=log)
        log(f"targets done: {data['sub'].shape[0]} fits, {data['failed']} failed, "
            f"{data['seconds']:.1f} s, mean EM iters {data['iters'].mean():.1f}")
        net = SetNet(a.d, a.K, seed=a.net_seed)
        tr = train(net, data, a.steps, batch=a.batch, lr=a.lr, gam=a.gam, log=log)
        if a.out:
            net.save(a.out)
            log(f"saved {a.out}")
    recs = evaluate(net, np.arange(a.eval_seed0, a.eval_seed0 + a.eval), a.eval_n or a.n, a.d, a.K, gen,
                    n_sub=a.keep, eps=a.eval_eps, g=a.g, log=log, chunk=a.eval_chunk,
                    bf16=not a.fp32_inference) if a.eval > 0 else []
    summ = summarise(recs) if recs else {}
    summ["inference"] = "fp32" if a.fp32_inference else "bf16 mfma"
    log(json.dumps(summ))
    if a.report:
        rep = {"args": vars(a), "train": tr, "summary": summ, "records": recs}
        if data is not None:
            rep.update({"targets_seconds": data["seconds"], "targets_failed": data["failed"],
                        "target_em_iters_mean": float(data["iters"].mean())})
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
