"""Every benchmark configuration at its real workload on the GPU, against the FP64 oracle.

  C2  1,000,000 x 128 mixture corpus, 10,000 queries, k in {10, 100}: the FP32 drop-in (heap and
      bitonic entry points, host buffers, as a C caller uses it) and the BF16 shortlist + exact
      rerank on a resident corpus;
  C3  10,000,000 x 96 mixture corpus, 10,000 queries, k = 10: BF16 and FP32 exact, FP8 E4M3 with
      recall@10 and distance error reported and bounded;
  C4  200,000 x 32 FP64 all-kNN excluding self, k = 8, the queries split into 64 uneven batches
      (log-uniform sizes in [64, 16384], seed 7) issued in a shuffled order as one packed search;
  C5  20,000,000 x 128 resident BF16 corpus, serving: requests of log-uniform size in [1, 4096]
      with k in {1, 10, 100} submitted together, each equal to the oracle on sampled rows and equal
      to the same request searched alone.

The oracle is the FP64 brute force with the canonical distance (oracle.exact_topk), run on a
sample of queries (every query of the sample: all corpus rows). Exact forms are asserted
bit-exact: identical indices, distances equal to the oracle's FP64 value rounded to the output
type.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"\n[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _sample(n, count, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=count, replace=False))


def _assert_exact(D, I, Do, Io, what):
    k = Io.shape[1]
    bad = np.nonzero(np.any(I[:, :k].astype(np.int64) != Io.astype(np.int64), axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} sampled queries differ from the oracle (first: {bad[:5]})"
    assert np.array_equal(D[:, :k], Do.astype(D.dtype)), f"{what}: distances differ from the oracle"


# ---------------------------------------------------------------------------------------------
# C2
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(oracle_mod):
    R, Q = oracle_mod.workload_c2()
    s = _sample(Q.shape[0], 128, 21)
    Do, Io = oracle_mod.exact_topk(Q[s], R, 100)
    return R, Q, s, Do, Io


@pytest.mark.timeout(600)
@pytest.mark.parametrize("entry", ["heap", "bitonic"])
def test_c2_dropin_fp32(gpu, knns_mod, c2, entry):
    R, Q, s, Do, Io = c2
    fn = knns_mod.gpuknns_heap if entry == "heap" else knns_mod.gpuknns_bitonic
    for k in (10, 100):
        _case("c2_dropin", entry=entry, k=k, m=R.shape[0], n=Q.shape[0])
        vals, inds, _ = fn(Q, R, k)
        _assert_exact(vals[s], inds[s], Do[:, :k], Io[:, :k], f"C2 drop-in {entry} k={k}")


@pytest.mark.timeout(600)
def test_c2_bf16_shortlist_rerank(gpu, knns_mod, c2):
    R, Q, s, Do, Io = c2
    with knns_mod.Corpus(R, bf16=True) as c:
        for k in (10, 100):
            for select in ("threshold", "bitonic"):
                _case("c2_bf16", k=k, select=select)
                D, I = c.search(Q, k, precision="bf16", select=select)
                _assert_exact(D[s], I[s], Do[:, :k], Io[:, :k], f"C2 BF16 k={k} {select}")
        # the resident FP32 form agrees on every query, not only the sample
        _case("c2_fp32_vs_bf16_all", k=100)
        Df, If = c.search(Q, 100, precision="fp32")
        Db, Ib = c.search(Q, 100, precision="bf16")
        assert np.array_equal(If, Ib) and np.array_equal(Df, Db)


# ---------------------------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3(oracle_mod):
    R, Q = oracle_mod.workload_c3()
    s = _sample(Q.shape[0], 64, 31)
    Do, Io = oracle_mod.exact_topk(Q[s], R, 10)
    return R, Q, s, Do, Io


@pytest.mark.timeout(900)
def test_c3_bf16_fp32_exact_fp8_recall(gpu, knns_mod, oracle_mod, c3):
    R, Q, s, Do, Io = c3
    _case("c3_corpus", m=R.shape[0], d=R.shape[1])
    c = knns_mod.Corpus(R, bf16=True, fp8=True)
    try:
        for prec in ("bf16", "fp32"):
            _case("c3_exact", precision=prec, k=10)
            D, I = c.search(Q, 10, precision=prec)
            _assert_exact(D[s], I[s], Do, Io, f"C3 {prec}")
        # FP8 is the approximate form: recall and error are reported against the oracle and
        # bounded (shortlist 4k+32 by default, and the configuration's 4k = 40)
        for shortlist, min_recall in ((0, 0.97), (40, 0.93)):
            _case("c3_fp8", shortlist=shortlist)
            D, I = c.search(Q, 10, precision="fp8", shortlist=shortlist)
            rec = oracle_mod.recall_at_k(I[s], Io)
            Dp = oracle_mod.pair_dist(Q[s], R, I[s])
            err = oracle_mod.max_rel_err(Dp, Do)
            print(f"[c3 fp8] shortlist={shortlist or '4k+32'} recall@10={rec:.4f} max_rel_err={err:.3e}", flush=True)
            assert rec >= min_recall
            # returned distances are the exact distances of the returned rows
            assert np.array_equal(D[s], Dp.astype(np.float32))
            # a missed neighbour is replaced by one at most a few percent farther
            assert err < 0.05
    finally:
        c.close()


# ---------------------------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------------------------
@pytest.mark.timeout(600)
def test_c4_fp64_all_knn_64_uneven_batches(gpu, knns_mod, oracle_mod):
    X = oracle_mod.workload_c4()
    m = X.shape[0]
    sizes = oracle_mod.log_uniform_sizes(64, 64, 16384, seed=7, total=m)
    assert sizes.sum() == m and sizes.min() >= 1
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    order = np.random.default_rng(7).permutation(64)          # batches issued out of row order
    Qp = np.concatenate([X[starts[b]:starts[b] + sizes[b]] for b in order])
    rows = np.concatenate([np.arange(starts[b], starts[b] + sizes[b]) for b in order])
    _case("c4_batches", batches=64, m=m, smallest=int(sizes.min()), largest=int(sizes.max()))
    with knns_mod.Corpus(X) as c:
        D, I, off = c.search_batches(Qp, sizes[order], self_rows=starts[order], precision="fp64", k=8)
    D = D.reshape(m, 8)
    I = I.reshape(m, 8)
    s = _sample(m, 512, 41)
    Do, Io = oracle_mod.exact_topk(Qp[s], X, 8, self_ids=rows[s])
    _assert_exact(D[s], I[s], Do, Io, "C4 all-kNN")
    assert not np.any(I == rows[:, None]), "a query returned its own row"
    # every batch is searched like the same rows alone: spot-check the smallest batch
    b = int(np.argmin(sizes))
    with knns_mod.Corpus(X) as c:
        Da, Ia = c.search(X[starts[b]:starts[b] + sizes[b]], 8, precision="fp64", exclude_self=True,
                          self_offset=int(starts[b]))
    pos = int(np.nonzero(order == b)[0][0])
    q0 = int(sizes[order][:pos].sum())
    assert np.array_equal(Ia, I[q0:q0 + sizes[b]]) and np.array_equal(Da, D[q0:q0 + sizes[b]])


# ---------------------------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_c5_serving_on_20m_corpus(gpu, knns_mod, oracle_mod):
    R, Qpool = oracle_mod.workload_c5()
    _case("c5_corpus", m=R.shape[0], d=R.shape[1])
    g = np.random.default_rng(9)
    reqs = []
    for _ in range(6):
        n = int(min(4096, max(1, round(np.exp(g.uniform(0.0, np.log(4096.0)))))))
        s0 = int(g.integers(0, Qpool.shape[0] - n + 1))
        reqs.append((Qpool[s0:s0 + n], int(g.choice([1, 10, 100]))))
    reqs.append((Qpool[:1], 100))                         # a one-query request at the largest k
    c = knns_mod.Corpus(R, bf16=True)
    try:
        with knns_mod.Server(c, precision="bf16", max_k=100, max_batch=4096) as srv:
            tickets = [srv.submit(q, k) for q, k in reqs]
            mixed = [srv.result(t) for t in tickets]
        for (q, k), (Dm, Im) in zip(reqs, mixed):
            _case("c5_request", n=q.shape[0], k=k)
            Da, Ia = c.search(q, k, precision="bf16")
            assert np.array_equal(Im, Ia) and np.array_equal(Dm, Da), "request differs alone vs in the mix"
            s = _sample(q.shape[0], min(4, q.shape[0]), q.shape[0])
            Do, Io = oracle_mod.exact_topk(q[s], R, k)
            _assert_exact(Dm[s], Im[s], Do, Io, f"C5 request n={q.shape[0]} k={k}")
    finally:
        c.close()
