"""Serving loop (knns_server_*): continuous batching over a resident corpus. Several lanes (each a
workspace, a stream and one captured HIP graph per batch-capacity bucket) run batches at once, so a
batch is packed, uploaded and launched while the previous one runs; requests larger than a batch
are split across batches, work beyond the free lanes waits in the queue.

Contract: a request's results are identical whether it is served alone or batched with others,
and identical to an eager knns_search of its queries (same indices, same distances). The GPU
tests replay every bucket graph many times with eager searches in between and compare every
replay, show device overlap between consecutive batches (their device intervals, from events),
and serve requests larger than the largest bucket. CPU tests cover argument checks that run before
any device work and the struct layouts."""
import ctypes

import numpy as np
import pytest


def test_server_struct_layouts(knns_mod):
    # as gcc lays out include/knns_native.h
    assert ctypes.sizeof(knns_mod.ServerConfig) == 32 and knns_mod.ServerConfig.lanes.offset == 24
    assert ctypes.sizeof(knns_mod.ServerStats) == 120 and knns_mod.ServerStats.researched.offset == 112


def test_server_config_rejected_before_device_work(knns_mod):
    L = knns_mod.native()
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(1)   # never dereferenced: the configuration is checked first
    for prec, max_k, max_batch, budget, lanes, want in [
            (0, 0, 64, 0, 0, knns_mod.KNNS_ERR_INVALID), (0, 1025, 64, 0, 0, knns_mod.KNNS_ERR_INVALID),
            (0, 10, 0, 0, 0, knns_mod.KNNS_ERR_INVALID), (0, 10, 64, -1, 0, knns_mod.KNNS_ERR_INVALID),
            (7, 10, 64, 0, 0, knns_mod.KNNS_ERR_INVALID), (0, 10, 64, 0, 3, knns_mod.KNNS_ERR_INVALID),
            (0, 10, 64, 0, -1, knns_mod.KNNS_ERR_INVALID)]:
        cfg = knns_mod.ServerConfig(prec, max_k, max_batch, budget, lanes)
        assert L.knns_server_create(ctypes.byref(h), fake, ctypes.byref(cfg)) == want
        assert not h.value
    cfg = knns_mod.ServerConfig(0, 10, 64, 0)
    assert L.knns_server_create(ctypes.byref(h), None, ctypes.byref(cfg)) == knns_mod.KNNS_ERR_INVALID


def test_server_null_handle_calls(knns_mod):
    L = knns_mod.native()
    t = ctypes.c_ulonglong()
    x = np.zeros(4, np.float32)
    assert L.knns_server_submit(None, x.ctypes.data, 0, 1, 1, x.ctypes.data, x.ctypes.data,
                                ctypes.byref(t)) == knns_mod.KNNS_ERR_INVALID
    assert L.knns_server_wait(None, 1, 0.0) == knns_mod.KNNS_ERR_INVALID
    s = knns_mod.ServerStats()
    assert L.knns_server_stats_get(None, ctypes.byref(s), 0) == knns_mod.KNNS_ERR_INVALID
    assert L.knns_server_destroy(None) == 0


# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def mixture_corpus(oracle_mod):
    R = oracle_mod.mixture(40000, 128, 64, 0.1, center_seed=31, point_seed=32)
    Q = oracle_mod.mixture(3000, 128, 64, 0.1, center_seed=31, point_seed=33)
    return R, Q


def _requests(Q, count, seed, kset=(1, 10, 100), nmax=200):
    g = np.random.default_rng(seed)
    reqs, pos = [], 0
    for _ in range(count):
        n = int(g.integers(1, nmax + 1))
        if pos + n > Q.shape[0]:
            pos = 0
        reqs.append((Q[pos:pos + n], int(g.choice(kset))))
        pos += n
    return reqs


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16", "fp32", "fp8"])
def test_serving_mix_equals_alone_equals_eager(gpu, knns_mod, mixture_corpus, precision):
    R, Q = mixture_corpus
    reqs = _requests(Q, 40, seed=7)
    print(f"\n[case] serving precision={precision} m={R.shape[0]} requests={len(reqs)}", flush=True)
    with knns_mod.Corpus(R, bf16=precision == "bf16", fp8=precision == "fp8") as corpus:
        eager = [corpus.search(q, k, precision=precision) for q, k in reqs]
        with knns_mod.Server(corpus, precision=precision, max_k=100, max_batch=1024) as srv:
            # 1. all requests in flight together (batched into mixes)
            tickets = [srv.submit(q, k) for q, k in reqs]
            mixed = [srv.result(t) for t in tickets]
            st = srv.stats()
            assert st["requests"] == len(reqs) and st["batches"] < len(reqs)
            # 2. each request alone, an eager search between replays
            alone = []
            for (q, k), (De, Ie) in zip(reqs, eager):
                alone.append(srv.result(srv.submit(q, k)))
                D2, I2 = corpus.search(q, k, precision=precision)
                assert np.array_equal(I2, Ie) and np.array_equal(D2, De)
            st2 = srv.stats()
            assert st2["batches"] >= st["batches"] + len(reqs)
            assert st2["p50_ms"] > 0 and st2["p99_ms"] >= st2["p50_ms"]
    for (De, Ie), (Dm, Im), (Da, Ia) in zip(eager, mixed, alone):
        assert np.array_equal(Im, Ie) and np.array_equal(Dm, De)
        assert np.array_equal(Ia, Ie) and np.array_equal(Da, De)


@pytest.mark.gpu
def test_serving_fp64(gpu, knns_mod, oracle_mod):
    R = oracle_mod.uniform(5000, 24, seed=41, lo=-1.0, hi=1.0, dtype=np.float64)
    Q = oracle_mod.uniform(300, 24, seed=42, lo=-1.0, hi=1.0, dtype=np.float64)
    reqs = _requests(Q, 12, seed=3, kset=(1, 8, 33), nmax=60)
    print("\n[case] serving fp64", flush=True)
    with knns_mod.Corpus(R) as corpus, knns_mod.Server(corpus, precision="fp64", max_k=40, max_batch=256) as srv:
        tickets = [srv.submit(q, k) for q, k in reqs]
        for (q, k), t in zip(reqs, tickets):
            D, I = srv.result(t)
            Do, Io = oracle_mod.exact_topk(q, R, k)
            assert np.array_equal(I, Io) and np.array_equal(D, Do)


@pytest.mark.gpu
def test_serving_admission(gpu, knns_mod, mixture_corpus):
    R, Q = mixture_corpus
    print("\n[case] serving admission", flush=True)
    with knns_mod.Corpus(R, bf16=True) as corpus:
        with knns_mod.Server(corpus, precision="bf16", max_k=100, max_batch=4096) as big:
            full = big.stats()
        assert full["buckets"] == 4 and full["max_batch"] == 4096   # 64, 256, 1024, 4096
        # a budget that holds the smaller buckets only: capacity shrinks, and a request larger than
        # the largest remaining bucket is deferred and served in several batches, not refused
        budget = full["workspace_bytes"] // 3
        with knns_mod.Server(corpus, precision="bf16", max_k=100, max_batch=4096, mem_budget=budget) as srv:
            st = srv.stats()
            assert 1 <= st["buckets"] < 4 and st["workspace_bytes"] <= budget
            with pytest.raises(knns_mod.KnnsError):
                srv.submit(Q[:4], 101)   # k above max_k
            big = Q[:2 * st["max_batch"] + 5]
            D, I = srv.result(srv.submit(big, 10))
            De, Ie = corpus.search(big, 10, precision="bf16")
            assert np.array_equal(I, Ie) and np.array_equal(D, De)
            assert srv.stats()["batches"] >= 3
        with pytest.raises(knns_mod.KnnsError):
            knns_mod.Server(corpus, precision="bf16", max_k=100, max_batch=4096, mem_budget=1000)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [2, 1])
def test_batches_overlap_on_the_device(gpu, knns_mod, oracle_mod, lanes):
    """Requests submitted while earlier batches run: with two lanes the next batch is launched
    before the previous one finished (its device interval starts before the previous ends); with
    one lane batches run one after another. Results are the eager search's either way."""
    R = oracle_mod.mixture(400000, 128, 256, 0.1, center_seed=51, point_seed=52)
    Q = oracle_mod.mixture(6000, 128, 256, 0.1, center_seed=51, point_seed=53)
    reqs = [(Q[i * 500:(i + 1) * 500], (10, 100, 1)[i % 3]) for i in range(12)]
    print(f"\n[case] serving overlap lanes={lanes}", flush=True)
    with knns_mod.Corpus(R, bf16=True) as corpus:
        eager = [corpus.search(q, k, precision="bf16") for q, k in reqs]
        with knns_mod.Server(corpus, precision="bf16", max_k=100, max_batch=1024, lanes=lanes) as srv:
            srv.result(srv.submit(Q[:64], 10))      # warm
            srv.stats(reset=True)
            tickets = []
            for q, k in reqs:                        # arrivals spread over the first batches
                tickets.append(srv.submit(q, k))
            got = [srv.result(t) for t in tickets]
            st = srv.stats()
    assert st["lanes"] == lanes and st["batches"] >= 6
    if lanes == 2:
        assert st["overlapped"] >= 1
    else:
        assert st["overlapped"] == 0
    for (De, Ie), (D, I) in zip(eager, got):
        assert np.array_equal(I, Ie) and np.array_equal(D, De)


@pytest.mark.gpu
def test_result_timeout_keeps_the_request(gpu, knns_mod, mixture_corpus):
    """A wait that times out leaves the request live with its output buffers: waiting again
    returns the full result."""
    R, Q = mixture_corpus
    print("\n[case] serving timeout", flush=True)
    with knns_mod.Corpus(R, bf16=True) as corpus:
        De, Ie = corpus.search(Q, 100, precision="bf16")
        with knns_mod.Server(corpus, precision="bf16", max_k=100, max_batch=1024) as srv:
            t = srv.submit(Q, 100)
            try:
                srv.result(t, timeout=0.0)
            except TimeoutError:
                pass
            D, I = srv.result(t)
            with pytest.raises(knns_mod.KnnsError):
                srv.result(t)                        # a done ticket is gone
    assert np.array_equal(I, Ie) and np.array_equal(D, De)
