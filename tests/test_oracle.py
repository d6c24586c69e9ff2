"""Pins the CPU oracle (oracle/) before anything is compared against it.

The reference ships no golden vectors (its demos draw unseeded rand data, demo/knnTest.m), so the
oracle is pinned by what the mathematics fixes: exact integer arithmetic, closed forms, brute
force on tiny inputs, invariances, and agreement between independent restatements (the knn.m
transcription, the FP64 brute force, and Python restatements of the reference's TBiS / heap /
host-merge selection run on the reference's own key matrix).
"""
import itertools

import numpy as np
import pytest


def _int_data(rng, n, d, lo=-3, hi=4):
    return rng.integers(lo, hi, size=(n, d)).astype(np.float32)


def _py_exact(Q, R, k, self_ids=None):
    """Pure-Python exact integer distances, (dist, idx) order."""
    out_d, out_i = [], []
    for i, q in enumerate(Q.astype(np.int64).tolist()):
        cand = []
        for j, r in enumerate(R.astype(np.int64).tolist()):
            if self_ids is not None and self_ids[i] == j:
                continue
            cand.append((sum((a - b) * (a - b) for a, b in zip(q, r)), j))
        cand.sort()
        out_d.append([c[0] for c in cand[:k]])
        out_i.append([c[1] for c in cand[:k]])
    return np.array(out_d, dtype=np.float64), np.array(out_i, dtype=np.int32)


def test_exact_topk_integer_closed_form(oracle_mod):
    rng = np.random.default_rng(5)
    R = _int_data(rng, 300, 7)
    R[50:60] = R[10]           # duplicated rows: ties must go to the lower index
    Q = _int_data(rng, 20, 7)
    Q[3] = R[10]
    for k in (1, 5, 37):
        D, I = oracle_mod.exact_topk(Q, R, k)
        Dp, Ip = _py_exact(Q, R, k)
        assert np.array_equal(I, Ip)
        assert np.array_equal(D, Dp)


def test_exact_topk_self_exclusion(oracle_mod):
    rng = np.random.default_rng(6)
    R = _int_data(rng, 120, 5)
    self_ids = np.arange(120, dtype=np.int64)
    D, I = oracle_mod.exact_topk(R, R, 8, self_ids=self_ids)
    Dp, Ip = _py_exact(R, R, 8, self_ids=self_ids)
    assert np.array_equal(I, Ip) and np.array_equal(D, Dp)
    assert not np.any(I == np.arange(120)[:, None])


def test_exact_topk_full_sort_and_float_brute(oracle_mod):
    R = oracle_mod.uniform(257, 43, 11)
    Q = oracle_mod.uniform(9, 43, 12)
    D, I = oracle_mod.exact_topk(Q, R, 257)          # k = m: a full sort
    dd = ((Q.astype(np.float64)[:, None, :] - R.astype(np.float64)[None]) ** 2).sum(-1)
    assert np.allclose(np.take_along_axis(dd, I, 1), D, rtol=1e-13, atol=0)
    assert np.all(np.diff(D, axis=1) >= 0)
    # fma-sequential vs numpy pairwise sums agree to rounding; the order agrees except at
    # near-ties closer than that rounding
    Ib = np.argsort(dd, axis=1, kind="stable")
    gap = np.abs(np.diff(np.sort(dd, axis=1), axis=1)).min()
    if gap > 1e-12:
        assert np.array_equal(I, Ib)


def test_exact_topk_invariances(oracle_mod):
    R = oracle_mod.uniform(500, 16, 21)
    Q = oracle_mod.uniform(12, 16, 22)
    D, I = oracle_mod.exact_topk(Q, R, 10)
    # scaling by a power of two scales every squared distance by its square, exactly
    D2, I2 = oracle_mod.exact_topk(Q * 8, R * 8, 10)
    assert np.array_equal(I, I2) and np.array_equal(D2, D * 64)
    # corpus permutation permutes the indices
    perm = np.random.default_rng(1).permutation(500)
    D3, I3 = oracle_mod.exact_topk(Q, R[perm], 10)
    assert np.array_equal(perm[I3], I) and np.array_equal(D3, D)
    # translation invariance (exact for small integers)
    rng = np.random.default_rng(2)
    Ri, Qi = _int_data(rng, 200, 6), _int_data(rng, 10, 6)
    a, b = oracle_mod.exact_topk(Qi, Ri, 7)
    c, e = oracle_mod.exact_topk(Qi + 5, Ri + 5, 7)
    assert np.array_equal(a, c) and np.array_equal(b, e)


def test_knnm_transcription_integer_exact(oracle_mod):
    """demo/knn.m in FP32 is exact on small integers, so it must equal the exact oracle,
    including the lower-index tie rule of MATLAB's stable sort (knn.m:5)."""
    rng = np.random.default_rng(7)
    R = _int_data(rng, 400, 9)
    R[100:110] = R[3]
    Q = _int_data(rng, 25, 9)
    for k in (1, 16, 64):
        Dm, Im = oracle_mod.knnm(Q, R, k)
        D, I = oracle_mod.exact_topk(Q, R, k)
        assert np.array_equal(Im, I)
        assert np.array_equal(Dm.astype(np.float64), D)
    Dd, Id = oracle_mod.knnm(Q.astype(np.float64), R.astype(np.float64), 16)
    D, I = oracle_mod.exact_topk(Q, R, 16)
    assert np.array_equal(Id, I) and np.array_equal(Dd, D)


def test_knnm_is_stable_sort_of_reference_keys(oracle_mod):
    R = oracle_mod.uniform(3000, 128, 0)
    Q = oracle_mod.uniform(8, 128, 1)
    K = oracle_mod.keys_f32(Q, R)
    Dm, Im = oracle_mod.knnm(Q, R, 16)
    Is = np.argsort(K, axis=1, kind="stable")[:, :16]
    assert np.array_equal(Im, Is)
    qn = np.array([np.float32(sum(np.float32(v) * np.float32(v) for v in q)) for q in Q])
    assert np.array_equal(Dm, np.take_along_axis(K, Is, 1) + qn[:, None])


def test_knnm_close_to_exact_on_c1_slice(oracle_mod):
    R, Q = oracle_mod.workload_c1()
    Q = Q[:32]
    Dm, Im = oracle_mod.knnm(Q, R, 16)
    D, I = oracle_mod.exact_topk(Q, R, 16)
    assert oracle_mod.max_rel_err(Dm, D) < 1e-4
    assert oracle_mod.recall_at_k(Im, I) > 0.99


@pytest.mark.parametrize("k", [1, 2, 16, 128])
@pytest.mark.parametrize("m", [100, 1024, 1500, 2500])
def test_ref_tbis_equals_sorted_topk(oracle_mod, k, m):
    if k > m:
        pytest.skip("k > m")
    rng = np.random.default_rng(m * 7 + k)
    keys = rng.permutation(m).astype(np.float32) * np.float32(0.5) + np.float32(1.0)
    tk, ti = oracle_mod.ref_tbis_select(keys, k)
    s = np.argsort(keys, kind="stable")[:k]
    assert np.array_equal(ti.astype(np.int64), s)
    assert np.array_equal(tk, keys[s])


@pytest.mark.parametrize("k", [1, 4, 16, 64])
def test_ref_heap_equals_topk_set(oracle_mod, k):
    rng = np.random.default_rng(k)
    keys = rng.permutation(3000).astype(np.float32)
    hk, hi = oracle_mod.ref_heap_select(keys, k)
    s = np.argsort(keys, kind="stable")[:k]
    assert sorted(hi.astype(np.int64).tolist()) == sorted(s.tolist())
    # heap order: slot 0 holds the maximum (cuknnHeap.cu:146-174)
    assert hk[0] == hk.max()


def test_ref_heap_ties_keep_first(oracle_mod):
    """Strict admission (cuknnHeap.cu:242): among equal keys the earliest rows stay."""
    keys = np.full(50, 3.0, dtype=np.float32)
    hk, hi = oracle_mod.ref_heap_select(keys, 8)
    assert sorted(hi.astype(int).tolist()) == list(range(8))


def test_ref_heap_non_power_of_two_k_is_truncated(oracle_mod):
    """qk = floor(log2 k) + 1 (cuknnHeap.cu:465): k=10 keeps only 8 live slots."""
    keys = np.arange(100, dtype=np.float32)[::-1].copy()
    hk, hi = oracle_mod.ref_heap_select(keys, 10)
    assert len(hk) == 8


def test_ref_merge_res_equals_union_topk(oracle_mod):
    rng = np.random.default_rng(3)
    P, Q, k = 4, 5, 8
    keys = rng.permutation(P * Q * k).astype(np.float32).reshape(P, Q, k)
    ids = np.arange(P * Q * k, dtype=np.float32).reshape(P, Q, k)
    md, mi = oracle_mod.ref_merge_res(keys.reshape(-1), ids.reshape(-1), k, Q, P)
    for q in range(Q):
        allk = keys[:, q, :].reshape(-1)
        alli = ids[:, q, :].reshape(-1)
        s = np.argsort(allk)[:k]
        assert sorted(mi[q].tolist()) == sorted(alli[s].tolist())


def test_reference_pipeline_equals_knnm(oracle_mod):
    """TBiS on the reference key matrix, then + ||q||^2 (dot3, utils.h:485-491) reproduces the
    knn.m transcription on tie-free inputs: both restate the same selection."""
    R = oracle_mod.uniform(2048 + 300, 32, 8)
    Q = oracle_mod.uniform(6, 32, 9)
    K = oracle_mod.keys_f32(Q, R)
    Dm, Im = oracle_mod.knnm(Q, R, 16)
    for i in range(Q.shape[0]):
        tk, ti = oracle_mod.ref_tbis_select(K[i], 16)
        assert np.array_equal(ti.astype(np.int32), Im[i])


def test_generators_deterministic_and_shaped(oracle_mod):
    a = oracle_mod.uniform(10, 4, 0)
    b = oracle_mod.uniform(10, 4, 0)
    assert a.dtype == np.float32 and np.array_equal(a, b) and a.min() >= 0 and a.max() < 1
    c = oracle_mod.uniform(10, 4, 6, lo=-1.0, hi=1.0, dtype=np.float64)
    assert c.dtype == np.float64 and c.min() >= -1 and c.max() < 1
    X = oracle_mod.mixture(20000, 16, 64, 0.1, center_seed=2, point_seed=3)
    C = oracle_mod.mixture_centers(64, 16, 2)
    d2 = ((X[:, None, :].astype(np.float64) - C[None]) ** 2).sum(-1).min(1)
    # every point sits within a few sigma of its centre: E||noise||^2 = d * sigma^2
    assert abs(d2.mean() - 16 * 0.01) < 0.01
    s = oracle_mod.log_uniform_sizes(64, 64, 16384, 7, total=200_000)
    assert s.sum() == 200_000 and len(s) == 64 and s.min() >= 1


def test_exact_topk_and_knnm_property_tiny_integer_inputs(oracle_mod):
    """Random tiny shapes with many exact ties (hypothesis): the FP64 brute force equals the pure
    Python integer brute force, and the knn.m transcription (exact on small integers) agrees."""
    hypothesis = pytest.importorskip("hypothesis")
    from hypothesis import given, settings, strategies as hst

    @settings(max_examples=60, deadline=None, derandomize=True)
    @given(m=hst.integers(1, 40), n=hst.integers(1, 6), d=hst.integers(1, 6), k=hst.integers(1, 45),
           hi=hst.integers(1, 3), seed=hst.integers(0, 2 ** 16), excl=hst.booleans())
    def check(m, n, d, k, hi, seed, excl):
        rng = np.random.default_rng(seed)
        R = rng.integers(0, hi + 1, (m, d)).astype(np.float32)
        Q = R[:n].copy() if excl and n <= m else rng.integers(0, hi + 1, (n, d)).astype(np.float32)
        self_ids = np.arange(n) if excl and n <= m else None
        kk = min(k, m - (1 if self_ids is not None else 0))
        if kk < 1:
            return
        D, I = oracle_mod.exact_topk(Q, R, kk, self_ids=self_ids)
        Dp, Ip = _py_exact(Q, R, kk, self_ids=self_ids)
        assert np.array_equal(I, Ip) and np.array_equal(D, Dp)
        if self_ids is None:
            Dm, Im = oracle_mod.knnm(Q, R, kk)
            assert np.array_equal(Im, Ip) and np.array_equal(Dm.astype(np.float64), Dp)

    check()


def _bucket_brute(Q, R, k, bucket_sizes, query_offsets, query_sizes):
    """Independent restatement for bucketed_topk: exact integer distances to EVERY corpus row, rows
    outside the query's bucket masked out by a row-id range test, then a stable sort over the whole
    corpus (so the indices are corpus rows without any per-bucket offset arithmetic)."""
    n, m = Q.shape[0], R.shape[0]
    D = np.zeros((n, k))
    I = np.zeros((n, k), dtype=np.int64)
    d2 = ((Q.astype(np.int64)[:, None, :] - R.astype(np.int64)[None, :, :]) ** 2).sum(-1)
    ends = np.cumsum(bucket_sizes)
    rows = np.arange(m)
    for j, (q0, nq) in enumerate(zip(query_offsets, query_sizes)):
        lo, hi = ends[j] - bucket_sizes[j], ends[j]
        inb = (rows >= lo) & (rows < hi)
        for i in range(q0, q0 + nq):
            key = np.where(inb, d2[i], np.iinfo(np.int64).max)
            order = np.lexsort((rows, key))[:k]
            ok = inb[order]
            D[i] = np.where(ok, key[order].astype(np.float64), np.inf)
            I[i] = np.where(ok, order, -1)
    return D, I


def test_bucketed_topk_pinned(oracle_mod):
    rng = np.random.default_rng(17)
    R = _int_data(rng, 230, 6)
    R[100:110] = R[3]                     # ties inside bucket 1 and across buckets 0 / 1
    Q = _int_data(rng, 40, 6)
    k = 9
    # one bucket holding the corpus, one cluster holding every query: exact_topk itself
    D1, I1 = oracle_mod.bucketed_topk(Q, R, k, [230], [0], [40])
    De, Ie = oracle_mod.exact_topk(Q, R, k)
    assert np.array_equal(I1, Ie) and np.array_equal(D1, De)
    # uneven buckets, a bucket smaller than k, an empty bucket without queries, a cluster without
    # queries, clusters whose query rows are not in cluster order, rows past the last bucket
    bs, qo, qs = [60, 90, 5, 0, 70], [30, 0, 12, 40, 20], [10, 12, 8, 0, 10]
    D, I = oracle_mod.bucketed_topk(Q, R, k, bs, qo, qs)
    Db, Ib = _bucket_brute(Q, R, k, bs, qo, qs)
    assert np.array_equal(I, Ib) and np.array_equal(D, Db)
    assert np.all(I[12:20, 5:] == -1) and np.all(I[12:20, :5] >= 150) and np.all(I[12:20, :5] < 155)
    assert np.all((I[20:30] >= 155) & (I[20:30] < 225))


# ---- pins of the helpers the GPU tests assert through: pair_dist, canonical_dist_np, recall_at_k,
# max_rel_err (each checked against values computed independently of the helper itself) ----
def test_pair_dist_matches_integer_brute_force_on_shuffled_candidates(oracle_mod):
    rng = np.random.default_rng(41)
    n, m, d, c = 7, 90, 11, 5          # n, m, d, c all different: a transposed index fails
    R = _int_data(rng, m, d)
    Q = _int_data(rng, n, d)
    C = np.stack([rng.permutation(m)[:c] for _ in range(n)]).astype(np.int32)
    D = oracle_mod.pair_dist(Q, R, C)
    Qi, Ri = Q.astype(np.int64).tolist(), R.astype(np.int64).tolist()
    want = [[sum((a - b) ** 2 for a, b in zip(Qi[i], Ri[C[i, t]])) for t in range(c)] for i in range(n)]
    assert np.array_equal(D, np.array(want, dtype=np.float64))
    # the same pairs in double precision, and the pairs exact_topk ranks first
    assert np.array_equal(oracle_mod.pair_dist(Q.astype(np.float64), R.astype(np.float64), C), D)
    Dt, It = oracle_mod.exact_topk(Q, R, c)
    assert np.array_equal(oracle_mod.pair_dist(Q, R, It), Dt)


def test_pair_dist_float_pairs_equal_the_python_canonical_form(oracle_mod):
    rng = np.random.default_rng(42)
    R = oracle_mod.uniform(64, 19, 43)
    Q = oracle_mod.uniform(3, 19, 44)
    C = np.array([[5, 63, 0, 5], [17, 2, 40, 41], [1, 1, 60, 33]], dtype=np.int32)
    D = oracle_mod.pair_dist(Q, R, C)
    for i in range(3):
        for t in range(4):
            assert D[i, t] == oracle_mod.canonical_dist_np(Q[i], R[C[i, t]])


def test_pair_dist_empty_slots_and_bad_arguments(oracle_mod):
    R = np.arange(12, dtype=np.float32).reshape(4, 3)
    Q = np.zeros((2, 3), dtype=np.float32)
    D = oracle_mod.pair_dist(Q, R, np.array([[-1, 0], [3, -1]], dtype=np.int32))
    assert np.isinf(D[0, 0]) and np.isinf(D[1, 1])
    assert D[0, 1] == 0 + 1 + 4 and D[1, 0] == 81 + 100 + 121
    with pytest.raises(ValueError):
        oracle_mod.pair_dist(Q, R, np.array([[4, 0], [0, 0]], dtype=np.int32))
    with pytest.raises(TypeError):
        oracle_mod.pair_dist(Q, R.astype(np.float64), np.zeros((2, 1), dtype=np.int32))


def test_canonical_dist_np_is_the_sequential_fma_chain():
    import math
    from fractions import Fraction

    import oracle

    # integers: the exact sum
    q = np.array([3, -2, 7, 0], dtype=np.float32)
    r = np.array([-1, 5, 7, 2], dtype=np.float32)
    assert oracle.canonical_dist_np(q, r) == 16 + 49 + 0 + 4
    # floats: fma keeps every product exact, so the result is the exact rational sum rounded at
    # each step; within d ulps of the exact value, and equal to an explicit left-to-right fold
    rng = np.random.default_rng(45)
    q = rng.random(33).astype(np.float32)
    r = rng.random(33).astype(np.float32)
    exact = sum((Fraction(float(a)) - Fraction(float(b))) ** 2 for a, b in zip(q, r))
    got = oracle.canonical_dist_np(q, r)
    assert abs(Fraction(got) - exact) <= 33 * 2.0 ** -52 * exact
    acc = Fraction(0)
    for a, b in zip(q.tolist(), r.tolist()):             # round after every step, exactly
        t = Fraction(float(a) - float(b))
        acc = Fraction(float(t * t + acc))
    assert got == float(acc)
    # a plain multiply-then-add (two roundings, the form this helper fell back to before) is a
    # different expression: on random pairs it differs from the fma chain somewhere
    differs = 0
    for s in range(50):
        q2, r2 = rng.random(33).astype(np.float32), rng.random(33).astype(np.float32)
        plain = 0.0
        for a, b in zip(q2.tolist(), r2.tolist()):
            plain = plain + (a - b) * (a - b)
        differs += plain != oracle.canonical_dist_np(q2, r2)
    assert differs > 0
    assert math.isfinite(got)


def test_recall_at_k_hand_computed():
    import oracle

    true = np.array([[0, 1, 2, 3], [10, 11, 12, 13]])
    assert oracle.recall_at_k(true.copy(), true) == 1.0
    got = np.array([[3, 2, 9, 8], [13, 99, 98, 97]])           # 2/4 and 1/4 (order does not matter)
    assert oracle.recall_at_k(got, true) == 3 / 8
    dup = np.array([[0, 0, 0, 1], [-1, -1, 10, 11]])           # a repeat counts once; -1 never
    assert oracle.recall_at_k(dup, true) == 4 / 8
    wide = np.array([[7, 0, 1, 2, 3], [10, 11, 12, 99, 13]])   # only the first k columns count
    assert oracle.recall_at_k(wide, true) == 6 / 8


def test_max_rel_err_hand_computed():
    import oracle

    t = np.array([[1.0, 2.0], [4.0, 0.0]])
    assert oracle.max_rel_err(t.copy(), t) == 0.0
    g = np.array([[1.5, 2.0], [3.0, 0.0]])                     # 0.5 and 0.25 relative
    assert oracle.max_rel_err(g, t) == 0.5
    g0 = np.array([[1.0, 2.0], [4.0, 1e-300]])                 # a zero distance compares absolutely
    assert oracle.max_rel_err(g0, t) == pytest.approx(1e-300 / np.finfo(np.float64).tiny)
    assert oracle.max_rel_err(np.zeros((0, 3)), np.zeros((0, 3))) == 0.0
    with pytest.raises(ValueError):
        oracle.max_rel_err(np.zeros((2, 2)), np.zeros((2, 3)))
