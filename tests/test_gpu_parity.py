"""GPU parity of the HIP path against the CPU oracle (runs on the MI355X box: -m gpu).

Contract (FP32 drop-in): indices identical to the FP64 oracle (exact canonical distance, ties by
lower index) and distances equal to the oracle's FP64 distance rounded to the output type -- the
rerank evaluates the same FMA sequence, so this is bit-exact. Against the reference semantics
(demo/knn.m ranking fl(||r||^2) - 2 fl(q.r) in FP32) indices may differ only between candidates
whose reference keys are closer than that model's own rounding error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(oracle_mod):
    R, Q = oracle_mod.workload_c1()
    D, I = oracle_mod.exact_topk(Q, R, 16)
    return R, Q, D, I


def _print_case(name, **kw):
    print(f"\n[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("entry", ["bitonic", "heap"])
def test_c1_dropin_matches_fp64_oracle(gpu, knns_mod, c1, entry):
    R, Q, D, I = c1
    _print_case("c1_dropin", entry=entry)
    fn = knns_mod.gpuknns_bitonic if entry == "bitonic" else knns_mod.gpuknns_heap
    vals, inds, t = fn(Q, R, 16)
    assert t > 0
    assert np.array_equal(inds.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(vals, D.astype(np.float32))


def test_c1_dropin_vs_reference_semantics(gpu, knns_mod, oracle_mod, c1):
    """Tier (ii): vs the knn.m transcription, set differences only between candidates whose
    exact keys ||r||^2 - 2 q.r are closer than the FP32 model's worst-case rounding
    d*u*(||r||^2 + 2|q.r|) of the two keys; distances within the model's error."""
    R, Q, D, I = c1
    _print_case("c1_vs_knnm")
    Dm, Im = oracle_mod.knnm(Q, R, 16)
    vals, inds, _ = knns_mod.gpuknns_heap(Q, R, 16)
    inds = inds.astype(np.int64)
    u = np.finfo(np.float32).eps / 2
    d = Q.shape[1]
    rn = (R.astype(np.float64) ** 2).sum(1)
    for i in range(Q.shape[0]):
        A = np.setdiff1d(inds[i], Im[i])
        B = np.setdiff1d(Im[i], inds[i])
        if len(A) == 0:
            continue
        q = Q[i].astype(np.float64)
        key = lambda j: rn[j] - 2 * (R[j].astype(np.float64) @ q)
        eps = lambda j: d * u * (rn[j] + 2 * np.abs(R[j].astype(np.float64) @ q))
        for a in A:
            for b in B:
                assert key(b) - key(a) <= eps(a) + eps(b), (i, a, b)
    assert np.max(np.abs(vals.astype(np.float64) - Dm) / Dm) < 1e-4


def test_double_library_bit_exact(gpu, knns_mod, oracle_mod):
    R = oracle_mod.uniform(8192, 50, 30, dtype=np.float64)
    Q = oracle_mod.uniform(300, 50, 31, dtype=np.float64)
    _print_case("double_dropin", m=8192, n=300, d=50, k=16)
    D, I = oracle_mod.exact_topk(Q, R, 16)
    for fn in (knns_mod.gpuknns_bitonic, knns_mod.gpuknns_heap):
        vals, inds, _ = fn(Q, R, 16, double=True)
        assert vals.dtype == np.float64
        assert np.array_equal(inds.astype(np.int64), I.astype(np.int64))
        assert np.array_equal(vals, D)


@pytest.mark.parametrize("m,n,d,k", [
    (1, 1, 1, 1), (5, 3, 3, 5), (63, 7, 2, 4), (64, 33, 17, 16), (100, 65, 43, 32),
    (1000, 1, 128, 1), (1000, 50, 128, 1000), (3000, 129, 200, 64), (2500, 40, 1024, 16),
    (1500, 20, 2048, 128), (5000, 70, 96, 100), (4097, 31, 32, 10)])
def test_shapes_match_oracle(gpu, knns_mod, oracle_mod, m, n, d, k):
    _print_case("shapes", m=m, n=n, d=d, k=k)
    R = oracle_mod.uniform(m, d, 100 + m)
    Q = oracle_mod.uniform(n, d, 200 + n)
    D, I = oracle_mod.exact_topk(Q, R, min(k, m))
    vals, inds, _ = knns_mod.gpuknns_heap(Q, R, k)
    kk = min(k, m)
    assert np.array_equal(inds[:, :kk].astype(np.int64), I.astype(np.int64))
    assert np.array_equal(vals[:, :kk], D.astype(np.float32))
    if k > m:
        assert np.all(inds[:, m:] == -1)


@pytest.mark.parametrize("double", [False, True])
def test_bucketed_search_matches_oracle(gpu, knns_mod, oracle_mod, double):
    """gpuknnLSH (include/cuknns.h:79): each query cluster's exact top-k within its bucket, with
    corpus-row indices, equal to oracle.bucketed_topk -- indices identical and distances equal to
    the FP64 oracle's rounded to knntype (the gpuknnsHeap contract). Covers a bucket smaller than
    k (+inf / -1 slots), an empty bucket, a cluster without queries, clusters whose query rows are
    out of cluster order, query rows in no cluster (left untouched) and corpus rows past the last
    bucket."""
    dt = np.float64 if double else np.float32
    R = oracle_mod.uniform(20000, 64, 41, dtype=dt)
    Q = oracle_mod.uniform(700, 64, 42, dtype=dt)
    bs, qo, qs = [7000, 5, 0, 9000, 3000], [300, 0, 250, 250, 600], [300, 200, 0, 50, 100]
    k = 32
    _print_case("bucketed", double=double, m=20000, n=700, d=64, k=k, clusters=len(bs))
    D, I = oracle_mod.bucketed_topk(Q, R, k, bs, qo, qs)
    vals, inds, t = knns_mod.gpuknn_lsh(Q, R, k, bs, qo, qs, double=double)
    assert t > 0 and vals.dtype == dt
    assert np.array_equal(inds.astype(np.int64), I)
    assert np.array_equal(vals, D.astype(dt))
    assert np.all(inds[0:200, 5:] == -1) and np.all(np.isinf(vals[0:200, 5:]))
    assert np.all(vals[200:250] == 0)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8", "fp64"])
def test_row_range_search_matches_oracle(gpu, knns_mod, oracle_mod, prec):
    """knns_search_rows: the top-k over corpus rows [lo, hi) only, with corpus-row indices, equal to
    the oracle over that slice, on integer data |x| <= 15 (exact keys in every format). Ranges start
    off the scan's 64/128-row stage grid, one ends at the last corpus row (its last stage reads into
    the corpus' slack rows), one holds a single row, one holds fewer rows than k (+inf / -1 slots),
    one is the whole corpus, and duplicated rows straddle a range start (ties by lower index)."""
    m, d, n, k = 9000, 40, 150, 24
    g = np.random.default_rng(23)
    dt = np.float64 if prec == "fp64" else np.float32
    R = g.integers(-15, 16, (m, d)).astype(dt)
    R[3000:3010] = R[2995]
    Q = g.integers(-15, 16, (n, d)).astype(dt)
    Q[:5] = R[2995]
    _print_case("row_range", prec=prec, m=m, d=d, n=n, k=k)
    with knns_mod.Corpus(R, bf16=prec == "bf16", fp8=prec == "fp8") as c:
        for lo, hi in [(2997, 7421), (8833, 9000), (1, 9000), (4444, 4445), (77, 90), (0, 9000), (129, 1153)]:
            kk = min(k, hi - lo)
            D, I = oracle_mod.exact_topk(Q, R[lo:hi], kk)
            dd, ii = c.search_rows(Q, k, lo, hi, precision=prec)
            assert np.array_equal(ii[:, :kk], I.astype(np.int32) + lo), (lo, hi)
            assert np.array_equal(dd[:, :kk], D.astype(dd.dtype)), (lo, hi)
            assert np.all(ii[:, kk:] == -1) and np.all(np.isinf(dd[:, kk:]))
        with pytest.raises(knns_mod.KnnsError, match="row range"):
            c.search_rows(Q, k, 5, 9001, precision=prec)
        with pytest.raises(knns_mod.KnnsError, match="row range"):
            c.search_rows(Q, k, 10, 10, precision=prec)


@pytest.mark.parametrize("double", [False, True])
def test_corpus_upload_chunks(gpu, knns_mod, oracle_mod, double):
    """A host corpus goes up in ~64 MB chunks through two alternating pinned halves. With 172-byte
    (float, d=43) and 400-byte (double, d=50) rows the 400,000-row corpus spans 2 and 3 chunks whose
    boundaries fall at rows 390,167 and 167,772 / 335,544: a query planted on a row of every chunk
    (and next to each boundary) must find it, and sampled queries match the oracle exactly."""
    dt = np.float64 if double else np.float32
    m, d, k = 400_000, 50 if double else 43, 8
    row_bytes = d * (8 if double else 4)
    chunk = (64 << 20) // row_bytes
    g = np.random.default_rng(31)
    R = g.integers(-20, 21, (m, d)).astype(dt)
    rows = sorted({1, chunk - 1, chunk, chunk + 1, min(2 * chunk + 5, m - 1), m - 1, 123_457})
    Q = np.concatenate([R[rows] + dt(0.25), g.integers(-20, 21, (40, d)).astype(dt)])
    _print_case("upload_chunks", double=double, m=m, d=d, chunk=chunk, planted=rows)
    vals, inds, _ = knns_mod.gpuknns_heap(Q, R, k, double=double)
    assert [int(inds[i, 0]) for i in range(len(rows))] == rows
    D, I = oracle_mod.exact_topk(Q, R, k)
    assert np.array_equal(inds.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(vals, D.astype(dt))


def test_ties_go_to_lower_index(gpu, knns_mod, oracle_mod):
    rng = np.random.default_rng(3)
    base = rng.integers(-4, 5, size=(700, 12)).astype(np.float32)
    R = np.concatenate([base, base, base[:300]])        # every row duplicated
    R = R[rng.permutation(len(R))]
    Q = rng.integers(-4, 5, size=(90, 12)).astype(np.float32)
    _print_case("ties", m=len(R), n=90, d=12, k=24)
    D, I = oracle_mod.exact_topk(Q, R, 24)
    for fn in (knns_mod.gpuknns_heap, knns_mod.gpuknns_bitonic):
        vals, inds, _ = fn(Q, R, 24)
        assert np.array_equal(inds.astype(np.int64), I.astype(np.int64))
        assert np.array_equal(vals, D.astype(np.float32))


def test_native_precisions_agree(gpu, knns_mod, oracle_mod):
    """BF16 shortlist + exact rerank returns the FP32 form's answer (recall 1.0 vs FP64)."""
    R = oracle_mod.mixture(60_000, 128, 256, 0.1, center_seed=2, point_seed=1002)
    Q = oracle_mod.mixture(300, 128, 256, 0.1, center_seed=2, point_seed=1003)
    _print_case("precisions", m=60000, n=300, d=128, k=100)
    D, I = oracle_mod.exact_topk(Q, R, 100)
    with knns_mod.Corpus(R, bf16=True) as c:
        for prec in ("fp32", "bf16"):
            for sel in ("threshold", "bitonic"):
                d, i = c.search(Q, 100, precision=prec, select=sel)
                assert np.array_equal(i.astype(np.int64), I.astype(np.int64)), (prec, sel)
                assert np.array_equal(d, D.astype(np.float32))


def test_fp64_all_knn_exclude_self(gpu, knns_mod, oracle_mod):
    P = oracle_mod.uniform(20_000, 32, 6, lo=-1.0, hi=1.0, dtype=np.float64)
    _print_case("allknn_fp64", m=20000, d=32, k=8)
    Qs = P[:1000]
    D, I = oracle_mod.exact_topk(Qs, P, 8, self_ids=np.arange(1000))
    with knns_mod.Corpus(P) as c:
        d, i = c.search(Qs, 8, precision="fp64", exclude_self=True, self_offset=0)
    assert np.array_equal(i.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(d, D)


def test_mixed_k_batch_equals_individual(gpu, knns_mod, oracle_mod):
    R = oracle_mod.uniform(20_000, 64, 40)
    Q = oracle_mod.uniform(200, 64, 41)
    kq = np.random.default_rng(0).choice([1, 10, 100], size=200).astype(np.int32)
    _print_case("mixed_k", m=20000, n=200)
    with knns_mod.Corpus(R) as c:
        res = c.search_multi(Q, kq)
        for i in range(0, 200, 37):
            d1, i1 = c.search(Q[i:i + 1], int(kq[i]))
            assert np.array_equal(res[i][1], i1[0]) and np.array_equal(res[i][0], d1[0])
    D, I = oracle_mod.exact_topk(Q, R, 100)
    for i in range(200):
        assert np.array_equal(res[i][1], I[i, :kq[i]])


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_adversarial_order_forces_in_scan_compaction(gpu, knns_mod, oracle_mod, prec):
    """Corpus rows get closer to every query as the row index grows, so every row beats the
    running threshold and the per-lane sub-pools overflow inside a round: the in-kernel exact
    K-th compaction must run and keep the right candidates."""
    rng = np.random.default_rng(11)
    m, d, n, k = 120_000, 64, 70, 20
    # row j sits at distance (m - j) from the origin along axis 0; queries sit at the origin
    R = np.zeros((m, d), dtype=np.float32)
    R[:, 0] = np.arange(m, 0, -1, dtype=np.float32)
    Q = (rng.standard_normal((n, d)) * 1e-3).astype(np.float32)
    _print_case("adversarial", m=m, n=n, d=d, k=k, prec=prec)
    D, I = oracle_mod.exact_topk(Q, R, k)
    with knns_mod.Corpus(R, bf16=(prec == "bf16")) as c:
        dd, ii = c.search(Q, k, precision=prec, slices=2)   # few slices: sub-pools must overflow
        st = knns_mod.last_stats()
    assert st["slow_compactions"] > 0
    assert np.array_equal(ii.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(dd, D.astype(np.float32))


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
@pytest.mark.parametrize("m,n,d,k", [
    (1, 1, 1, 1), (5, 3, 3, 9), (100, 1, 17, 100), (300, 2, 40, 7), (3000, 33, 130, 10),
    (20_000, 50, 64, 1024), (4097, 70, 96, 500), (2000, 129, 256, 64), (2000, 40, 1024, 16)])
def test_low_precision_shapes_on_representable_data(gpu, knns_mod, oracle_mod, prec, m, n, d, k):
    """BF16 and FP8 on integer data |x| <= 15 (exact in both formats, so the shortlist keys are the
    exact keys): every shape returns the exact answer. Covers one-query waves, k > m (padding:
    index -1, distance +inf), and shortlists above the merge's register capacity (k = 1024 gives
    a BF16 shortlist of 1552, FP8 2048: the merge's global-memory select)."""
    if prec == "fp8" and d > 256:
        pytest.skip("FP8 rows are built for d <= 256 (wider corpora are rejected: tests/test_fp8.py)")
    _print_case("lowprec_shapes", prec=prec, m=m, n=n, d=d, k=k)
    g = np.random.default_rng(m * 7 + d)
    R = g.integers(-15, 16, (m, d)).astype(np.float32)
    Q = g.integers(-15, 16, (n, d)).astype(np.float32)
    kk = min(k, m)
    D, I = oracle_mod.exact_topk(Q, R, kk)
    with knns_mod.Corpus(R, bf16=prec == "bf16", fp8=prec == "fp8") as c:
        dd, ii = c.search(Q, k, precision=prec)
    assert np.array_equal(ii[:, :kk].astype(np.int64), I.astype(np.int64))
    assert np.array_equal(dd[:, :kk], D.astype(np.float32))
    if k > m:
        assert np.all(ii[:, m:] == -1) and np.all(np.isinf(dd[:, m:]))


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_paths_on_representable_data(gpu, knns_mod, oracle_mod, prec):
    """The native API's other paths, per precision, on integer data (exact keys in every format):
    all-kNN with self exclusion (including duplicated rows), per-query k, device pointers on a
    caller's stream, double output distances for a float corpus."""
    import torch

    _print_case("paths", prec=prec)
    g = np.random.default_rng(17)
    m, d = 6000, 40
    R = g.integers(-15, 16, (m, d)).astype(np.float32)
    R[100:110] = R[7]                      # duplicates: ties go to the lower index, self by index
    with knns_mod.Corpus(R, bf16=prec == "bf16", fp8=prec == "fp8") as c:
        # all-kNN over the first 500 rows, self excluded by index
        D, I = oracle_mod.exact_topk(R[:500], R, 12, self_ids=np.arange(500))
        dd, ii = c.search(R[:500], 12, precision=prec, exclude_self=True, self_offset=0)
        assert np.array_equal(ii.astype(np.int64), I.astype(np.int64))
        assert np.array_equal(dd, D.astype(np.float32))
        # per-query k
        Q = g.integers(-15, 16, (90, d)).astype(np.float32)
        kq = g.choice([1, 5, 40, 130], size=90).astype(np.int32)
        res = c.search_multi(Q, kq, precision=prec)
        D, I = oracle_mod.exact_topk(Q, R, 130)
        for i in range(90):
            assert np.array_equal(res[i][1].astype(np.int64), I[i, :kq[i]].astype(np.int64)), i
            assert np.array_equal(res[i][0], D[i, :kq[i]].astype(np.float32)), i
        # device pointers on a caller's stream
        dev = torch.device("cuda:0")
        s = torch.cuda.Stream(device=dev)
        Qd = torch.from_numpy(Q).to(dev)
        Dd = torch.empty((90, 33), dtype=torch.float32, device=dev)
        Id = torch.empty((90, 33), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        c.search_device(Qd, 90, 33, Dd, Id, precision=prec, stream=s.cuda_stream)
        assert np.array_equal(Id.cpu().numpy().astype(np.int64), I[:, :33].astype(np.int64))
        assert np.array_equal(Dd.cpu().numpy(), D[:, :33].astype(np.float32))
        # double distances for a float corpus: the FP64 oracle's values exactly
        d64, i64 = c.search(Q, 20, precision=prec, dist_f64=True)
        assert d64.dtype == np.float64
        assert np.array_equal(i64.astype(np.int64), I[:, :20].astype(np.int64))
        assert np.array_equal(d64, D[:, :20])


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_metamorphic_invariants(gpu, knns_mod, oracle_mod, prec):
    """Properties the exact answer has whatever the oracle says, on continuous data (no exact
    distance ties): (1) permuting the corpus rows permutes the returned indices and leaves the
    distances unchanged; (2) scaling queries and corpus by 2^j scales every distance by 4^j
    exactly (BF16 and FP32 keys, the FP64 rerank and the final rounding all commute with a power
    of two) and leaves the indices unchanged; (3) k=1 returns the argmin of the FP64 distances;
    (4) k=m returns the whole corpus sorted by (distance, index)."""
    _print_case("metamorphic", prec=prec)
    g = np.random.default_rng(23)
    m, n, d, k = 3000, 70, 40, 16

    def search(Q, R, kk):
        if prec == "fp32":
            dist, idx, _ = knns_mod.gpuknns_heap(Q, R, kk)
            return dist, idx.astype(np.int64)
        with knns_mod.Corpus(R, bf16=True) as c:
            dist, idx = c.search(Q, kk, precision="bf16")
        return dist, idx.astype(np.int64)

    R = g.standard_normal((m, d)).astype(np.float32)
    Q = g.standard_normal((n, d)).astype(np.float32)
    D0, I0 = search(Q, R, k)
    Do, Io = oracle_mod.exact_topk(Q, R, k)
    assert np.array_equal(I0, Io.astype(np.int64)) and np.array_equal(D0, Do.astype(np.float32))
    # (1) permutation of the corpus
    p = g.permutation(m)
    D1, I1 = search(Q, R[p], k)
    assert np.array_equal(p[I1], I0) and np.array_equal(D1, D0)
    # (2) scaling by powers of two (no overflow or underflow at these magnitudes)
    for j in (3, -5):
        s = np.float32(2.0 ** j)
        D2, I2 = search(Q * s, R * s, k)
        assert np.array_equal(I2, I0), j
        assert np.array_equal(D2, D0 * np.float32(4.0 ** j)), j
    # (3) k = 1 is the argmin of the FP64 distances
    full = ((Q.astype(np.float64)[:, None, :] - R.astype(np.float64)[None, :, :]) ** 2).sum(-1)
    D3, I3 = search(Q, R, 1)
    assert np.array_equal(I3[:, 0], np.argmin(full, axis=1))
    # (4) k = m on a small corpus: every row, sorted by (distance, index)
    Rs = R[:200]
    D4, I4 = search(Q[:10], Rs, 200)
    Df, If = oracle_mod.exact_topk(Q[:10], Rs, 200)
    assert np.array_equal(I4, If.astype(np.int64)) and np.array_equal(D4, Df.astype(np.float32))
    assert all(sorted(r.tolist()) == list(range(200)) for r in I4)


def test_large_corpus_row_offsets(gpu, knns_mod, oracle_mod):
    """A corpus of 20,000,003 rows (not a multiple of the 128-row stage): row indices above 2^24
    come back exactly (int32 through the native API) and the scans address stages whose byte
    offsets pass 2 GiB (FP32 operand rows: 2.9 GB). Integer data (exact in BF16) makes the BF16
    answer the exact one too; each query's nearest rows are planted above row 2^24."""
    m, d, n, k = 20_000_003, 32, 8, 10
    _print_case("large_corpus", m=m, d=d, n=n, k=k)
    g = np.random.default_rng(29)
    R = g.integers(0, 16, (m, d), dtype=np.int8).astype(np.float32)
    Q = g.integers(0, 16, (n, d), dtype=np.int8).astype(np.float32)
    for i in range(n):
        R[(1 << 24) + 1 + 7 * i] = Q[i]            # distance 0
        R[m - 1 - i] = Q[i]
        R[m - 1 - i, i] += 1.0                      # distance 1
    D, I = oracle_mod.exact_topk(Q, R, k)
    assert np.all(I[:, 0] > (1 << 24)) and np.all(I[:, 1] >= m - n)
    with knns_mod.Corpus(R, bf16=True) as c:
        for prec in ("fp32", "bf16"):
            dd, ii = c.search(Q, k, precision=prec)
            assert np.array_equal(ii.astype(np.int64), I.astype(np.int64)), prec
            assert np.array_equal(dd, D.astype(np.float32)), prec


@pytest.mark.gpu
def test_dropin_stage_times_are_opt_in(gpu, knns_mod, oracle_mod):
    """Per-kernel device events cost time, so the drop-in records stage times only on request
    (knnsSetStageTimes); the whole-call time is always there, and results do not change."""
    R = oracle_mod.uniform(20000, 64, 71)
    Q = oracle_mod.uniform(200, 64, 72)
    try:
        knns_mod.dropin_stage_times(False)
        v0, i0, _ = knns_mod.gpuknns_heap(Q, R, 10)
        t0 = knns_mod.dropin_last_times()
        assert t0["knn_time"] > 0 and t0["dst_time"] == 0 and t0["srch_time"] == 0
        knns_mod.dropin_stage_times(True)
        v1, i1, _ = knns_mod.gpuknns_heap(Q, R, 10)
        t1 = knns_mod.dropin_last_times()
        assert t1["dst_time"] > 0 and t1["srch_time"] > 0 and t1["knn_time"] > 0
    finally:
        knns_mod.dropin_stage_times(False)
    assert np.array_equal(i0, i1) and np.array_equal(v0, v1)
