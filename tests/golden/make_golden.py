"""Writes the golden fixtures of tests/test_golden.py (run from the repository root):

    python tests/golden/make_golden.py

The reference ships no vectors (data/data.txt is a date stub), so these are ours. They are made
with numpy alone -- not with oracle/ -- so that they check the oracle as well as the GPU path.
All inputs are small integers: every squared distance is an exact integer in FP64 (and in FP32
below 2^24, and in BF16 operands for values <= 256), so the expected top-k is unambiguous
whatever the summation order: ascending distance, ties to the lower corpus index (the stable
sort of demo/knn.m:5). Each case stores corpus, queries, k and the expected (dist, idx)."""
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def topk(Q, R, k, self_offset=None):
    Qf, Rf = Q.astype(np.float64), R.astype(np.float64)
    D = ((Qf[:, None, :] - Rf[None, :, :]) ** 2).sum(axis=2)
    if self_offset is not None:
        for i in range(Q.shape[0]):
            D[i, self_offset + i] = np.inf
    I = np.argsort(D, axis=1, kind="stable")[:, :k]
    Dk = np.take_along_axis(D, I, axis=1)
    if k > R.shape[0]:   # fewer rows than k: padded with +inf / -1
        pad = k - R.shape[0]
        Dk = np.concatenate([Dk, np.full((Q.shape[0], pad), np.inf)], axis=1)
        I = np.concatenate([I, np.full((Q.shape[0], pad), -1)], axis=1)
    return Dk, I.astype(np.int32)


def case(name, R, Q, k, dtype, self_offset=None):
    R = R.astype(dtype)
    Q = Q.astype(dtype)
    D, I = topk(Q, R, k, self_offset)
    kw = dict(corpus=R, queries=Q, k=np.int32(k), dist=D, idx=I)
    if self_offset is not None:
        kw["self_offset"] = np.int32(self_offset)
    np.savez_compressed(OUT / f"{name}.npz", **kw)
    print(f"{name}: corpus {R.shape} {R.dtype}, queries {Q.shape}, k={k}")


def main():
    g = np.random.Generator(np.random.PCG64(20240601))
    # random small integers: few exact ties
    case("int_uniform_f32", g.integers(-50, 51, (700, 19)), g.integers(-50, 51, (23, 19)), 10, np.float32)
    # duplicated corpus rows and a coarse grid: many exact ties, resolved to the lower index
    base = g.integers(0, 4, (60, 6))
    R = np.concatenate([base, base[::-1], base[:17]])
    case("int_ties_f32", R, g.integers(0, 4, (15, 6)), 24, np.float32)
    # k larger than the corpus: padded with +inf / -1
    case("int_small_corpus_f32", g.integers(-9, 10, (5, 3)), g.integers(-9, 10, (4, 3)), 8, np.float32)
    # double precision
    case("int_uniform_f64", g.integers(-1000, 1001, (400, 33)), g.integers(-1000, 1001, (9, 33)), 16, np.float64)
    # all-kNN excluding self: queries are corpus rows 100..139
    R = g.integers(-20, 21, (300, 5))
    case("int_allknn_f64", R, R[100:140], 6, np.float64, self_offset=100)
    # values <= 256 are exact in BF16: the BF16 shortlist path must return the exact answer too
    case("int_bf16_exact_f32", g.integers(0, 257, (1000, 128)), g.integers(0, 257, (24, 128)), 50, np.float32)


if __name__ == "__main__":
    main()
