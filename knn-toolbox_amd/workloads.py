"""Synthetic workloads of the benchmark and the tests (numpy only; no device code).

Every generator is seeded (numpy PCG64), so a workload is the same array on every machine:
* ``uniform``      -- i.i.d. U[lo, hi), drawn in float64 and rounded to the requested type;
* ``mixture``      -- equal-weight isotropic Gaussian mixture (centres ~ N(0, I), spread sigma);
* ``log_uniform_sizes`` -- request / batch sizes log-uniform in [lo, hi];
* ``workload_c1`` .. ``workload_c5`` -- the corpora and query sets the benchmark runs:
    c1  16,384 x 128 U[0,1) corpus, 1,024 queries, k=16 (the reference demos' scale);
    c2  1,000,000 x 128 mixture corpus, 10,000 queries, k=100 (the headline);
    c3  10,000,000 x 96 mixture corpus, 10,000 queries, k=10 (BF16 / FP8 shortlists);
    c4  200,000 x 32 FP64 U[-1,1) set, all-kNN, k=8;
    c5  20,000,000 x 128 mixture corpus and a 20,000-query request pool (serving).
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, d: int, seed: int, lo: float = 0.0, hi: float = 1.0, dtype=np.float32) -> np.ndarray:
    """i.i.d. U[lo, hi) from numpy PCG64(seed), generated in float64 then rounded to dtype."""
    g = np.random.Generator(np.random.PCG64(seed))
    x = g.random((n, d), dtype=np.float64)
    return (lo + (hi - lo) * x).astype(dtype)


def mixture_centers(ncenters: int, d: int, seed: int) -> np.ndarray:
    g = np.random.Generator(np.random.PCG64(seed))
    return g.standard_normal((ncenters, d))


def mixture(n: int, d: int, ncenters: int, sigma: float, center_seed: int, point_seed: int,
            dtype=np.float32, chunk: int = 1 << 20) -> np.ndarray:
    """Equal-weight isotropic Gaussian mixture: centres ~ N(0, I) drawn from PCG64(center_seed);
    each point picks a centre uniformly and adds sigma * N(0, I), both from PCG64(point_seed),
    in chunks of `chunk` points (each chunk draws its centre picks, then its noise: the chunk size
    is part of the definition). Queries of a workload use the corpus' centres with their own
    point_seed."""
    C = mixture_centers(ncenters, d, center_seed)
    g = np.random.Generator(np.random.PCG64(point_seed))
    out = np.empty((n, d), dtype=dtype)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        a = g.integers(0, ncenters, size=e - s)
        out[s:e] = (C[a] + sigma * g.standard_normal((e - s, d))).astype(dtype)
    return out


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int, total: int | None = None) -> np.ndarray:
    """Batch sizes log-uniform in [lo, hi]; if total is given, rescaled to sum exactly to total by
    largest remainder."""
    g = np.random.Generator(np.random.PCG64(seed))
    s = np.exp(g.uniform(np.log(lo), np.log(hi), size=count))
    if total is None:
        return np.maximum(1, np.round(s)).astype(np.int64)
    w = s / s.sum() * total
    base = np.floor(w).astype(np.int64)
    rem = total - base.sum()
    order = np.argsort(-(w - base), kind="stable")
    base[order[:rem]] += 1
    return base


# Workload definitions. Each returns (corpus, queries), or the point set for all-kNN.
def workload_c1():
    """m=16384 x d=128 U[0,1) seed 0; n=1024 U[0,1) seed 1; k=16."""
    return uniform(16384, 128, 0), uniform(1024, 128, 1)


def workload_c2(m: int = 1_000_000, n: int = 10_000):
    """1024-centre mixture, sigma 0.1, d=128: corpus seed 2, queries seed 3 (same centres)."""
    R = mixture(m, 128, 1024, 0.1, center_seed=2, point_seed=2 + 1000)
    Q = mixture(n, 128, 1024, 0.1, center_seed=2, point_seed=3 + 1000)
    return R, Q


def workload_c5(m: int = 20_000_000, n: int = 20_000):
    """Serving corpus: 16384-centre mixture, sigma 0.1, d=128: corpus seed 8; the query pool that
    requests draw from uses the same centres (seed 10; seed 9 drives the arrival process)."""
    R = mixture(m, 128, 16384, 0.1, center_seed=8, point_seed=8 + 1000)
    Q = mixture(n, 128, 16384, 0.1, center_seed=8, point_seed=10 + 1000)
    return R, Q


def workload_c3(m: int = 10_000_000, n: int = 10_000):
    """4096-centre mixture, sigma 0.15, d=96: corpus seed 4, queries seed 5 (same centres)."""
    R = mixture(m, 96, 4096, 0.15, center_seed=4, point_seed=4 + 1000)
    Q = mixture(n, 96, 4096, 0.15, center_seed=4, point_seed=5 + 1000)
    return R, Q


def workload_c4(m: int = 200_000):
    """FP64 U[-1, 1) points, d=32, seed 6."""
    return uniform(m, 32, 6, lo=-1.0, hi=1.0, dtype=np.float64)
