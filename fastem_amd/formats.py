"""Readers and writers of the reference's on-disk formats.

.bin (Util/MultithreadedDataset.cpp:119-134): int32 N, int32 d; if N == -1
an int64 N follows; then N*d FP32 row-major.
.cmp (EM2/GaussianMixture.cpp:155-171,228-244): int32 K, int32 d, then per
component d*d FP64 covariance (row-major), d FP64 mean, FP64 weight; then a
trailing FP64 0.0.
"""
import struct

import numpy as np


def write_bin(path, X, long_header=False):
    """long_header forces the 16-byte header (-1, d, int64 N) that files with
    more than 2^31 - 1 rows need."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    with open(path, "wb") as f:
        if long_header or n > 2**31 - 1:
            f.write(struct.pack("<iiq", -1, d, n))
        else:
            f.write(struct.pack("<ii", n, d))
        f.write(X.tobytes())


def read_bin(path):
    with open(path, "rb") as f:
        n, d = struct.unpack("<ii", f.read(8))
        if n == -1:
            (n,) = struct.unpack("<q", f.read(8))
        X = np.frombuffer(f.read(n * d * 4), dtype=np.float32).reshape(n, d)
    return X


def write_cmp(path, w, mu, cov):
    w = np.asarray(w, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    cov = np.asarray(cov, dtype=np.float64)
    K, d = mu.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", K, d))
        for k in range(K):
            f.write(np.ascontiguousarray(cov[k]).tobytes())
            f.write(np.ascontiguousarray(mu[k]).tobytes())
            f.write(struct.pack("<d", w[k]))
        f.write(struct.pack("<d", 0.0))


def read_cmp(path):
    with open(path, "rb") as f:
        K, d = struct.unpack("<ii", f.read(8))
        w = np.empty(K)
        mu = np.empty((K, d))
        cov = np.empty((K, d, d))
        for k in range(K):
            cov[k] = np.frombuffer(f.read(8 * d * d), dtype=np.float64).reshape(d, d)
            mu[k] = np.frombuffer(f.read(8 * d), dtype=np.float64)
            (w[k],) = struct.unpack("<d", f.read(8))
    return w, mu, cov
