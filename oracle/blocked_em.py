"""Blocked FP64 restatement of the reference's SimpleMatrixEM on BLAS: the CPU
baseline.

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it and the tests
check it against the oracle (oracle.c); the product never imports it.

It keeps the reference's algorithm shape rather than the oracle's plain loops:
* the dataset is widened to an FP64 copy (EMAlgorithm::CopyDataset,
  EM2/EMAlgorithm.cpp:225-235);
* rows split into T contiguous thread partitions, each split into nblocks
  blocks (SimpleBlockManager, EM2/BlockManager.cpp:59-77, FindOpenMPItems,
  Util/OpenMP.h:37-46), nblocks from `-N automax:<L3 MB>`
  (EM2/BlockManager.cpp:10-28);
* E-step per block and component: T = X - mu_k, T <- T U_k by dtrmm (right,
  upper), s = -0.5 rowsum(T^2) + c_k, then the row log-sum-exp, posteriors and
  LL (EM2/MatrixEM.cpp:39-87,275-307);
* M-step per block: N_k += colsum(R), M += R^T X by dgemm; after the thread
  reduction mu = M / N_k; per block and component T = (X - mu_k) sqrt(r/N_k)
  and S_k += T^T T by dsyrk (upper); reduction; lower = upper^T; pi = N_k / N
  (EM2/MatrixEM.cpp:53-67,90-99,309-378).
The T "threads" are T worker processes over the partitions (the dataset and
posteriors in shared memory), each calling single-threaded BLAS (numpy's
OpenBLAS through scipy.linalg.blas), as the reference calls sequential MKL
inside OpenMP threads; Python threads would serialise on the interpreter lock
in the BLAS wrappers.  Partial sums are combined in thread order (the naive
reducer, Util/OMPReducer.h:51-56).
"""
import math
import os

import numpy as np
import scipy.linalg
from scipy.linalg import blas
from threadpoolctl import threadpool_limits


def find_items(n, size, ident):
    """Util/OpenMP.h:37-46 FindOpenMPItems -> (offset, count)."""
    c, rem = divmod(n, size)
    if ident < rem:
        c += 1
        return c * ident, c
    return c * ident + rem, c


def automax_blocks(n, d, K, l3_mb):
    """EM2/BlockManager.cpp:10-16: blocks so that a block's rows of X and of
    the posteriors (16 (d + K) bytes per row) fill L3 about 1/1.2 full."""
    return max(1, int(n * (16.0 * d + 16.0 * K) / (l3_mb * 1024.0 * 1024.0) * 1.2))


def factor(covs, w):
    """NormalDensity::Init + CalcLogCoeffs: U_k = (L_k^-1)^T (upper) and
    c_k = log w_k - d/2 log 2 pi - logdet_k / 2."""
    K, d, _ = covs.shape
    U = np.empty_like(covs)
    c = np.empty(K)
    for k in range(K):
        C = np.tril(covs[k]) + np.tril(covs[k], -1).T
        L = np.linalg.cholesky(C)
        U[k] = scipy.linalg.solve_triangular(L, np.eye(d), lower=True).T
        c[k] = math.log(w[k]) - 0.5 * d * math.log(2.0 * math.pi) - np.log(np.diag(L)).sum()
    return U, c


# ---- worker side: one process per "thread" partition ----------------------
_W = {}


def _attach(xname, rname, n, d, K, blocks):
    from multiprocessing import shared_memory
    threadpool_limits(1)  # sequential BLAS in each worker
    xs = shared_memory.SharedMemory(name=xname)
    rs = shared_memory.SharedMemory(name=rname)
    _W.update(xs=xs, rs=rs, X=np.ndarray((n, d), np.float64, xs.buf), R=np.ndarray((n, K), np.float64, rs.buf),
              K=K, d=d, blocks=blocks)


def _estep_part(t, mu, Ut, c):
    X, R, K = _W["X"], _W["R"], _W["K"]
    ll = 0.0
    for r0, cnt in _W["blocks"][t]:
        Xb = X[r0:r0 + cnt]
        S = np.empty((cnt, K))
        for k in range(K):
            # (X - mu_k) U_k, formed transposed: U_k^T (X - mu_k)^T with the
            # Fortran-ordered views BLAS takes without a copy
            Tt = (Xb - mu[k]).T
            Tt = blas.dtrmm(1.0, Ut[k], Tt, side=0, lower=1, overwrite_b=1)
            S[:, k] = -0.5 * np.einsum("ij,ij->j", Tt, Tt) + c[k]
        mx = S.max(axis=1, keepdims=True)
        lse = np.log(np.exp(S - mx).sum(axis=1)) + mx[:, 0]
        ll += float(np.add.reduce(lse))
        R[r0:r0 + cnt] = np.exp(S - lse[:, None])
    return ll


def _means_part(t):
    X, R, K, d = _W["X"], _W["R"], _W["K"], _W["d"]
    nk = np.zeros(K)
    M = np.zeros((K, d))
    for r0, cnt in _W["blocks"][t]:
        Rb, Xb = R[r0:r0 + cnt], X[r0:r0 + cnt]
        nk += Rb.sum(axis=0)
        M += Rb.T @ Xb  # dgemm
    return nk, M


def _cov_part(t, mu, nk):
    X, R, K, d = _W["X"], _W["R"], _W["K"], _W["d"]
    # S[:, :, k] is Fortran-contiguous: dsyrk updates it in place
    S = np.zeros((d, d, K), order="F")
    inv = 1.0 / np.sqrt(nk)
    for r0, cnt in _W["blocks"][t]:
        Rb, Xb = R[r0:r0 + cnt], X[r0:r0 + cnt]
        for k in range(K):
            Tt = ((Xb - mu[k]) * (np.sqrt(Rb[:, k]) * inv[k])[:, None]).T
            blas.dsyrk(1.0, Tt, beta=1.0, c=S[:, :, k], trans=0, lower=0, overwrite_c=1)
    return S


class BlockedEM:
    """One EM problem: X (n x d, FP32 in, widened to an FP64 copy in shared
    memory), `threads` worker processes, nblocks blocks per partition.  Use
    as a context manager (it owns shared memory and processes)."""

    def __init__(self, X, K, threads, nblocks):
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        from multiprocessing import shared_memory
        X = np.asarray(X)
        self.n, self.d = X.shape
        self.K, self.T, self.nb = K, threads, nblocks
        self._xs = shared_memory.SharedMemory(create=True, size=self.n * self.d * 8)
        self._rs = shared_memory.SharedMemory(create=True, size=self.n * K * 8)
        np.ndarray((self.n, self.d), np.float64, self._xs.buf)[:] = X
        blocks = []
        for t in range(threads):
            p0, pc = find_items(self.n, threads, t)
            blocks.append([(p0 + o, c) for o, c in (find_items(pc, nblocks, b) for b in range(nblocks)) if c > 0])
        # spawned workers: fresh interpreters that load numpy/scipy only
        self.pool = ProcessPoolExecutor(threads, mp_context=mp.get_context("spawn"), initializer=_attach,
                                        initargs=(self._xs.name, self._rs.name, self.n, self.d, K, blocks))

    def close(self):
        if self.pool is not None:
            self.pool.shutdown()
            self.pool = None
            for s in (self._xs, self._rs):
                s.close()
                s.unlink()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def posteriors(self):
        return np.ndarray((self.n, self.K), np.float64, self._rs.buf).copy()

    def iterate(self, w, mu, cov):
        """One EM iteration from (w, mu, cov): returns (LL, w', mu', cov')."""
        U, c = factor(cov, w)
        Ut = [np.asfortranarray(u.T) for u in U]
        T = range(self.T)
        lls = list(self.pool.map(_estep_part, T, [mu] * self.T, [Ut] * self.T, [c] * self.T))
        parts = list(self.pool.map(_means_part, T))
        nk = parts[0][0].copy()
        M = parts[0][1].copy()
        for a, b in parts[1:]:
            nk += a
            M += b
        mu2 = M / nk[:, None]
        cparts = list(self.pool.map(_cov_part, T, [mu2] * self.T, [nk] * self.T))
        S = cparts[0].copy()
        for p in cparts[1:]:
            S += p
        S = np.ascontiguousarray(np.transpose(S, (2, 0, 1)))
        cov2 = np.triu(S) + np.transpose(np.triu(S, 1), (0, 2, 1))
        ll = 0.0
        for v in lls:
            ll += v
        return ll, nk / self.n, mu2, cov2


def host_description():
    """nproc, CPU model, sockets and NUMA nodes of this host (Linux)."""
    model, sockets = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    try:
        numa = len([e for e in os.listdir("/sys/devices/system/node") if e.startswith("node") and e[4:].isdigit()])
    except OSError:
        numa = None
    l3 = None
    try:
        for idx in sorted(os.listdir("/sys/devices/system/cpu/cpu0/cache")):
            base = f"/sys/devices/system/cpu/cpu0/cache/{idx}"
            if open(f"{base}/level").read().strip() == "3":
                size = open(f"{base}/size").read().strip()
                l3 = float(size[:-1]) / 1024.0 if size.endswith("K") else float(size.rstrip("M"))
    except (OSError, ValueError):
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "sockets": len(sockets) or None, "numa_nodes": numa, "l3_mb_per_cpu0": l3}
