"""Shortlist certificate: exact top-k by construction for the FP32, BF16 and FP64 forms.

The scan ranks corpus rows by an approximate key s~ = ||r~||^2/2 - q~.r~ (operands rounded for
the MFMA, FP32 accumulation) and keeps a shortlist of K' rows; the final kernel reranks them
exactly. A query is certified when every row outside the shortlist provably lies farther than
the k-th exact distance D_k: with tau the K'-th key, alpha the accumulation bound and the rounding
radii rho_q = ||q~ - q||, rho_r = max ||r~ - r||,

    ||q - r|| >= sqrt(2 (tau - alpha) + ||q~||^2) - rho_q - rho_r  >  D_k.

Every threshold keeps the certificate band above its K'-th key (rows the bound cannot separate
from it), so the shortlist grows where neighbours are dense; a query whose band overflows the
carried shortlist is searched again with every row inside the bound admitted.

CPU: the bound is a true lower bound on vectors whose rounding errors are aligned against the
distance (the worst case of the triangle inequality), and leaving out either radius breaks it.
GPU: on a near-tie shell (333 rows per query within 1e-3 of the same squared distance) the BF16
shortlist of k + k/2 + 16 provably misses true neighbours without the certificate, while the
certified search (band kept in one pass) equals the FP64 oracle for FP32, BF16 and FP64, through
the native API, the packed batches, the no-allocation entry point and the drop-in; on a denser
shell (1,500 rows per query, beyond the carried shortlist) the certificate fails, the re-search
runs and the answer is again the oracle's.
"""
import numpy as np
import pytest


def _bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _key_f32(q, r):
    """s~ accumulated sequentially in FP32 from exact BF16 products (one of the orders the bound
    covers): -||r~||^2/2 first, then the products."""
    acc = np.float32(-0.5 * float((r.astype(np.float64) ** 2).sum()))
    for a, b in zip(q.astype(np.float32), r.astype(np.float32)):
        acc = np.float32(acc + np.float32(a * b))
    return -float(acc)


def test_bound_is_a_lower_bound_on_aligned_rounding():
    """Worst case of the triangle inequality: in every coordinate q rounds up and r rounds down by
    almost half a BF16 ulp while q - r points along the same positive direction, so
    ||q~ - r~|| - ||q - r|| is within 0.4% of rho_q + rho_r."""
    rng = np.random.default_rng(3)
    d = 64
    ulp = 2.0 ** -7                                   # BF16 ulp of values in [1, 2)
    ok_full, broken_q, broken_r = True, True, True
    for trial in range(20):
        b = (1.0 + rng.integers(8, 120, d) * ulp)    # representable BF16 values in [1, 2)
        step = rng.integers(1, 4) * ulp               # q - r along (1, ..., 1)
        q = (b + 0.501 * ulp).astype(np.float32)      # rounds up by 0.499 ulp
        r = (b - step + 0.499 * ulp).astype(np.float32)   # rounds down by 0.499 ulp
        qt, rt = _bf16(q).astype(np.float64), _bf16(r).astype(np.float64)
        q64, r64 = q.astype(np.float64), r.astype(np.float64)
        assert np.all(qt > q64) and np.all(rt < r64)
        true = np.sqrt(((q64 - r64) ** 2).sum())
        rho_q = np.sqrt(((qt - q64) ** 2).sum())
        rho_r = np.sqrt(((rt - r64) ** 2).sum())
        s = _key_f32(qt, rt)
        n = d + 16
        gamma = n * 2.0 ** -23 / (1 - n * 2.0 ** -23) + 2.0 ** -26
        nq, nr = np.sqrt((qt ** 2).sum()), np.sqrt((rt ** 2).sum())
        alpha = gamma * (nq * nr + 0.5 * nr * nr)
        inner = np.sqrt(max(0.0, 2 * (s - alpha) + (qt ** 2).sum()))
        ok_full &= bool(inner - rho_q - rho_r <= true)
        broken_q &= bool(inner - rho_r > true)       # a bound without the query's radius fails
        broken_r &= bool(inner - rho_q > true)       # ... and one without the corpus radius
    assert ok_full
    assert broken_q and broken_r


# ---------------------------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------------------------
def _shell(seed=5, d=128, n_shell=2000, n_far=30000, n_q=6):
    """Queries far from the origin (large BF16 rounding radii); around each, n_shell / n_q rows at
    squared distance 1 + eps (eps uniform in [0, 1e-3]) and far rows at squared distance >= 4."""
    g = np.random.default_rng(seed)
    Q = (g.standard_normal((n_q, d)) * 1.0).astype(np.float32)
    rows = []
    per = n_shell // n_q
    for i in range(n_q):
        u = g.standard_normal((per, d))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        rad = np.sqrt(1.0 + g.uniform(0.0, 1e-3, per))
        rows.append(Q[i].astype(np.float64) + rad[:, None] * u)
    u = g.standard_normal((n_far, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    far = Q[g.integers(0, n_q, n_far)].astype(np.float64) + (2.0 + g.uniform(0, 3, n_far))[:, None] * u
    R = np.concatenate(rows + [far]).astype(np.float32)
    R = R[g.permutation(R.shape[0])]
    return R, Q


def test_predicted_rank_bounds_the_binomial_tail(knns_mod):
    """Before the last round a certified search ranks by knns_predicted_rank(K', f): the number of a
    query's K' nearest rows inside a prefix holding the fraction f of rows in no particular order is
    Binomial(K', f), and the rank is an upper quantile of it (tail at most 1e-7), never above K',
    non-decreasing in f, and K' itself once every row is seen."""
    from scipy.stats import binom
    fs = [1e-6, 1e-4, 1e-3, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5, 0.7, 0.9, 0.99]
    for K in list(range(1, 301)) + list(range(308, 2049, 20)):
        prev = 0
        for f in fs:
            r = knns_mod.predicted_rank(K, f)
            assert prev <= r <= K
            prev = r
            if r < K:
                assert binom.sf(r - 1, K, f) <= 1e-7, (K, f, r)
        assert knns_mod.predicted_rank(K, 1.0) == K
    # the headline's early merges (K' = 118 after 1024, 4096, 16384 of 10^6 rows) rank far below K'
    assert all(knns_mod.predicted_rank(118, f) <= 20 for f in (0.001024, 0.004096, 0.016384))


@pytest.fixture(scope="module")
def shell(oracle_mod):
    R, Q = _shell()
    D, I = oracle_mod.exact_topk(Q, R, 100)
    return R, Q, D, I


def _case(name, **kw):
    print(f"\n[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_near_tie_shell_is_certified_exact(gpu, knns_mod, shell, prec):
    R, Q, D, I = shell
    with knns_mod.Corpus(R, bf16=True) as c:
        for k in (100, 10):
            _case("shell", prec=prec, k=k)
            Dg, Ig = c.search(Q, k, precision=prec)
            st = knns_mod.last_stats()
            assert np.array_equal(Ig, I[:, :k]) and np.array_equal(Dg, D[:, :k].astype(np.float32))
            assert st["researched"] == st["uncertified"]
            if prec == "bf16":
                # without the certificate the k + k/2 + 16 BF16 shortlist misses true neighbours
                Dn, In = c.search(Q, k, precision=prec, certify=False)
                assert knns_mod.last_stats()["researched"] == 0
                assert not np.array_equal(In, I[:, :k])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_band_overflow_is_researched(gpu, knns_mod, oracle_mod, prec):
    R, Q = _shell(seed=9, n_shell=3000, n_far=20000, n_q=2)      # 1,500 shell rows per query
    D, I = oracle_mod.exact_topk(Q, R, 100)
    _case("dense_shell", prec=prec)
    with knns_mod.Corpus(R, bf16=True) as c:
        Dg, Ig = c.search(Q, 100, precision=prec)
        st = knns_mod.last_stats()
    assert np.array_equal(Ig, I) and np.array_equal(Dg, D.astype(np.float32))
    assert st["uncertified"] > 0 and st["researched"] == st["uncertified"]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_prefix_holding_the_neighbours_is_researched(gpu, knns_mod, oracle_mod, prec):
    """Before the last round the merges rank by the predicted share of the K' nearest rows in the
    prefix seen so far (true for rows in no particular order). A corpus whose first rows are one
    query's 400 nearest breaks that prediction for the query: its shortlist comes out short, the
    certificate fails, the re-search runs and the answer is the oracle's; the other queries are
    certified in one pass."""
    R = oracle_mod.mixture(200000, 64, 64, 0.1, center_seed=91, point_seed=92)
    Q = oracle_mod.mixture(64, 64, 64, 0.1, center_seed=91, point_seed=93)
    d0 = ((R.astype(np.float64) - Q[0].astype(np.float64)) ** 2).sum(1)
    near = np.argsort(d0, kind="stable")[:400]
    rest = np.setdiff1d(np.arange(R.shape[0]), near)
    R = np.ascontiguousarray(R[np.concatenate([near, rest])])
    D, I = oracle_mod.exact_topk(Q, R, 100)
    _case("neighbours_first", prec=prec)
    with knns_mod.Corpus(R, bf16=True) as c:
        Dg, Ig = c.search(Q, 100, precision=prec)
        st = knns_mod.last_stats()
    assert np.array_equal(Ig, I) and np.array_equal(Dg, D.astype(np.float32))
    assert 1 <= st["researched"] < 16 and st["researched"] == st["uncertified"]


@pytest.mark.gpu
def test_certified_variants_on_the_shell(gpu, knns_mod, oracle_mod, shell):
    R, Q, D, I = shell
    with knns_mod.Corpus(R, bf16=True) as c:
        _case("shell_bitonic")
        Db, Ib = c.search(Q, 100, precision="bf16", select="bitonic")
        assert np.array_equal(Ib, I) and np.array_equal(Db, D.astype(np.float32))
        # packed batches with per-batch k and self exclusion: query rows are corpus rows here
        sizes = np.array([3, 2, 1])
        rows = np.array([10, 20, 30])
        Qr = np.concatenate([R[rows[b]:rows[b] + sizes[b]] for b in range(3)])
        rr = np.concatenate([np.arange(rows[b], rows[b] + sizes[b]) for b in range(3)])
        ks = np.array([100, 7, 50], dtype=np.int32)
        _case("shell_batches")
        Dp, Ip, off = c.search_batches(Qr, sizes, ks=ks, self_rows=rows, precision="bf16")
        kq = np.repeat(ks, sizes)
        Do, Io = oracle_mod.exact_topk(Qr, R, 100, self_ids=rr)
        for i in range(Qr.shape[0]):
            assert np.array_equal(Ip[off[i]:off[i + 1]], Io[i, :kq[i]])
            assert np.array_equal(Dp[off[i]:off[i + 1]], Do[i, :kq[i]].astype(np.float32))


@pytest.mark.gpu
def test_certified_fp64_and_dropin_on_the_shell(gpu, knns_mod, oracle_mod, shell):
    R, Q, D, I = shell
    _case("shell_dropin")
    vals, inds, _ = knns_mod.gpuknns_heap(Q, R, 100)
    assert np.array_equal(inds.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(vals, D.astype(np.float32))
    R64, Q64 = R.astype(np.float64), Q.astype(np.float64)
    D64, I64 = oracle_mod.exact_topk(Q64, R64, 100)
    _case("shell_fp64")
    with knns_mod.Corpus(R64) as c:
        Dg, Ig = c.search(Q64, 100, precision="fp64")
    assert np.array_equal(Ig, I64) and np.array_equal(Dg, D64)


@pytest.mark.gpu
def test_certified_noalloc_on_the_shell(gpu, knns_mod, shell):
    torch = pytest.importorskip("torch")
    R, Q, D, I = shell
    n, k = Q.shape[0], 100
    with knns_mod.Corpus(R, bf16=True) as c:
        _case("shell_noalloc")
        q = torch.from_numpy(Q).to("cuda:0")
        d = torch.empty((n, k), dtype=torch.float32, device="cuda:0")
        i = torch.empty((n, k), dtype=torch.int32, device="cuda:0")
        c.search_noalloc(q, n, k, d, i, precision="bf16")
        torch.cuda.synchronize()
    assert np.array_equal(i.cpu().numpy(), I) and np.array_equal(d.cpu().numpy(), D.astype(np.float32))


@pytest.mark.gpu
def test_easy_data_is_certified_in_one_pass(gpu, knns_mod, oracle_mod):
    R, Q = oracle_mod.workload_c1()
    with knns_mod.Corpus(R) as c:
        _case("c1_certified")
        Dg, Ig = c.search(Q, 16, precision="fp32")
        st = knns_mod.last_stats()
    assert st["uncertified"] == 0 and st["researched"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("call", ["dropin", "bf16"])
def test_bf16_unprovable_queries_fall_back_to_fp32(gpu, knns_mod, oracle_mod, call):
    """Rows 1 + 0.04 u: the BF16 rounding radius (up to 2^-8 per unit coordinate) is large against
    the neighbour spacing, so the BF16 error bound of every query admits 7,000-8,000 of the 8192
    rows (a NumPy model of the bound), more than the re-search's 2048-row shortlist; the FP32
    bound admits 200-450. Those queries are searched once more on the FP32 rows and the answer is
    the oracle's, through the drop-in call (which searches float inputs in BF16) and through a
    resident BF16 search."""
    g = np.random.default_rng(31)
    m, d, n, k = 8192, 32, 16, 10
    R = (1.0 + 0.04 * g.random((m, d))).astype(np.float32)
    Q = (1.0 + 0.04 * g.random((n, d))).astype(np.float32)
    D, I = oracle_mod.exact_topk(Q, R, k)
    _case(f"bf16_fallback_{call}", m=m, d=d, n=n, k=k)
    if call == "dropin":
        vals, inds, _ = knns_mod.gpuknns_heap(Q, R, k)
    else:
        with knns_mod.Corpus(R, bf16=True) as c:
            vals, inds = c.search(Q, k, precision="bf16")
    st = knns_mod.last_stats()
    assert st["researched"] == n
    assert np.array_equal(inds.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(vals, D.astype(np.float32))
