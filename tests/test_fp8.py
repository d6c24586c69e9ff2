"""FP8 (OCP E4M3, per-row power-of-two scale) shortlist + exact FP32 rerank.

The FP8 form ranks candidates by the key of the dequantised rows, ||q~ - r~||^2 with
r~ = 2^p * e4m3(r / 2^p), keeps a shortlist of c*k candidates and reranks them exactly, so
(1) with data that E4M3 represents exactly (integers |x| <= 15) its answer is the exact answer;
(2) with a shortlist of exactly k it returns the k best rows by the FP8 key (the oracle's
    fp8_keys), up to pairs whose keys are closer than the FP32 accumulation error;
(3) every returned distance is the exact distance of the returned row.
CPU tests pin the oracle's quantiser: against torch.float8_e4m3fn, the code table of the format,
and brute force for the scale exponent."""
import numpy as np
import pytest


# ---------------------------------------------------------------------------------------------
# CPU: the oracle's quantiser
# ---------------------------------------------------------------------------------------------
def test_e4m3_encode_matches_torch(oracle_mod):
    torch = pytest.importorskip("torch")
    g = np.random.default_rng(0)
    x = np.concatenate([g.uniform(-447.9, 447.9, 200_000), g.uniform(-0.02, 0.02, 100_000),
                        g.standard_normal(50_000)]).astype(np.float32)
    # every midpoint between neighbouring finite codes (ties to even), below the saturation
    vals = np.unique(oracle_mod.e4m3_decode(np.arange(0, 0x7F)))
    mids = ((vals[1:] + vals[:-1]) / 2).astype(np.float32)
    x = np.concatenate([x, mids, -mids, vals.astype(np.float32)])
    x = x[np.abs(x) < 448]
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(oracle_mod.e4m3_encode(x), want)


def test_e4m3_code_table(oracle_mod):
    enc = lambda v: int(oracle_mod.e4m3_encode(np.array([v]))[0])
    assert enc(1.0) == 0x38 and enc(-1.0) == 0xB8 and enc(-0.0) == 0x80 and enc(0.0) == 0x00
    assert enc(448.0) == 0x7E and enc(470.0) == 0x7E and enc(1e30) == 0x7E and enc(-1e30) == 0xFE
    assert enc(2.0 ** -9) == 0x01 and enc(2.0 ** -10) == 0x00 and enc(3 * 2.0 ** -10) == 0x02
    assert enc(2.0 ** -6) == 0x08 and enc(7 * 2.0 ** -9) == 0x07
    assert enc(17.0) == 0x58 and enc(18.0) == 0x59 and enc(19.0) == 0x5A   # 17, 19: ties to even
    # every finite code decodes and re-encodes to itself (0x80 is -0)
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F])
    assert np.array_equal(oracle_mod.e4m3_encode(oracle_mod.e4m3_decode(codes)), codes.astype(np.uint8))
    # decoded magnitudes increase with the code
    assert np.all(np.diff(oracle_mod.e4m3_decode(np.arange(0, 0x7F))) > 0)
    # integers up to 15 (4 significant bits) are exact
    ints = np.arange(-15, 16, dtype=np.float64)
    assert np.array_equal(oracle_mod.e4m3_decode(oracle_mod.e4m3_encode(ints)), ints)


def test_fp8_scale_exp_brute_force(oracle_mod):
    g = np.random.default_rng(1)
    amax = np.concatenate([np.ldexp(g.uniform(0.5, 1.0, 3000), g.integers(-100, 100, 3000)),
                           448.0 * np.ldexp(1.0, np.arange(-20, 20)), [0.0, 448.0, 449.0, 224.0]])
    got = oracle_mod.fp8_scale_exp(amax)
    for a, p in zip(amax, got):
        if a == 0:
            assert p == 0
            continue
        want = next(q for q in range(-126, 127) if a / 2.0 ** q <= 448.0)
        assert p == want, (a, p, want)


def test_fp8_rows_integer_rows_are_exact(oracle_mod):
    X = np.random.default_rng(2).integers(-15, 16, (50, 37)).astype(np.float32)
    codes, p, deq = oracle_mod.fp8_rows(X)
    assert np.array_equal(deq, X.astype(np.float64))
    assert np.all(np.abs(codes.astype(np.int64) & 0x7F) <= 0x7E)
    # the scale puts every row's largest magnitude into (224, 448]
    amax = np.abs(X).max(1)
    nz = amax > 0
    assert np.all((amax[nz] / 2.0 ** p[nz] > 224) & (amax[nz] / 2.0 ** p[nz] <= 448))


def test_fp8_keys_are_distances_of_dequantised_rows(oracle_mod):
    g = np.random.default_rng(3)
    Q = g.standard_normal((4, 9)).astype(np.float32)
    R = g.standard_normal((30, 9)).astype(np.float32)
    K = oracle_mod.fp8_keys(Q, R)
    _, _, Qd = oracle_mod.fp8_rows(Q)
    _, _, Rd = oracle_mod.fp8_rows(R)
    for i in range(4):
        for j in range(30):
            assert abs(K[i, j] - float(((Qd[i] - Rd[j]) ** 2).sum())) < 1e-12
    # the quantisation error is that of a 3-bit mantissa: |r~ - r| <= 2^-4 |r| (normals)
    big = np.abs(R) > 2.0 ** -3
    assert np.all(np.abs(Rd - R)[big] <= 2.0 ** -4 * np.abs(R)[big] + 1e-12)


def test_fp8_corpus_rejects_wide_rows_before_device_work(knns_mod):
    import ctypes

    L = knns_mod.native()
    h = ctypes.c_void_p()
    x = np.zeros((4, 300), np.float32)
    st = L.knns_corpus_create(ctypes.byref(h), x.ctypes.data, 0, 4, 300, knns_mod.KNNS_CORPUS_FP8)
    assert st == knns_mod.KNNS_ERR_UNSUPPORTED and not h.value


# ---------------------------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------------------------
def _case(name, **kw):
    print(f"\n[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("d", [24, 96, 200])          # FP8 rows of 64, 128 and 256 bytes
@pytest.mark.parametrize("n", [20, 150])              # one and two query blocks per wave
def test_fp8_exact_on_representable_data(gpu, knns_mod, oracle_mod, d, n):
    g = np.random.default_rng(d * 1000 + n)
    m, k = 5000, 10
    R = g.integers(-15, 16, (m, d)).astype(np.float32)
    Q = g.integers(-15, 16, (n, d)).astype(np.float32)
    _case("fp8_exact", m=m, n=n, d=d, k=k)
    D, I = oracle_mod.exact_topk(Q, R, k)
    with knns_mod.Corpus(R, fp8=True) as c:
        for sel in ("threshold", "bitonic"):
            dd, ii = c.search(Q, k, precision="fp8", select=sel)
            assert np.array_equal(ii.astype(np.int64), I.astype(np.int64)), sel
            assert np.array_equal(dd, D.astype(np.float32)), sel


@pytest.mark.gpu
@pytest.mark.parametrize("d", [64, 96, 256])
def test_fp8_shortlist_of_k_is_the_fp8_top_k(gpu, knns_mod, oracle_mod, d):
    g = np.random.default_rng(7 + d)
    m, n, k = 20_000, 96, 16
    R = g.standard_normal((m, d)).astype(np.float32)
    Q = g.standard_normal((n, d)).astype(np.float32)
    _case("fp8_keys", m=m, n=n, d=d, k=k)
    K = oracle_mod.fp8_keys(Q, R)
    with knns_mod.Corpus(R, fp8=True) as c:
        dd, ii = c.search(Q, k, precision="fp8", shortlist=k)
    # FP32 accumulation of the key: |error| well below 1e-5 of ||q~||^2 + ||r~||^2
    scale = (np.abs(K).max(axis=1) + 1.0)
    for i in range(n):
        order = np.argsort(K[i], kind="stable")
        kth = K[i, order[k - 1]]
        tol = 1e-5 * scale[i]
        got = set(ii[i].tolist())
        assert all(K[i, j] <= kth + tol for j in got), i
        assert all(j in got for j in order[:k] if K[i, j] < kth - tol), i
    # exact distances of the returned rows, sorted by (distance, index)
    Dp = oracle_mod.pair_dist(Q, R, ii)
    assert np.array_equal(dd, Dp.astype(np.float32))
    for i in range(n):
        assert all((dd[i, t], ii[i, t]) <= (dd[i, t + 1], ii[i, t + 1]) for t in range(k - 1))


@pytest.mark.gpu
def test_fp8_recall_on_clustered_data(gpu, knns_mod, oracle_mod):
    """Recall of the FP8 shortlist (default 4k+32) on a Gaussian mixture like the 10M x 96 one
    (4096 centres there; fewer points per centre here, so the same sigma is harder)."""
    R = oracle_mod.mixture(200_000, 96, 512, 0.15, center_seed=4, point_seed=1004)
    Q = oracle_mod.mixture(500, 96, 512, 0.15, center_seed=4, point_seed=1005)
    _case("fp8_recall", m=200000, n=500, d=96, k=10)
    D, I = oracle_mod.exact_topk(Q, R, 10)
    with knns_mod.Corpus(R, fp8=True) as c:
        dd, ii = c.search(Q, 10, precision="fp8")
        d40, i40 = c.search(Q, 10, precision="fp8", shortlist=40)
    rec = oracle_mod.recall_at_k(ii, I)
    rec40 = oracle_mod.recall_at_k(i40, I)
    print(f"fp8 recall@10: shortlist 72 {rec:.4f}, shortlist 40 {rec40:.4f}", flush=True)
    assert rec >= 0.98 and rec40 >= 0.95
    assert np.array_equal(dd, oracle_mod.pair_dist(Q, R, ii).astype(np.float32))


@pytest.mark.gpu
def test_fp8_all_knn_exclude_self_and_double_corpus(gpu, knns_mod, oracle_mod):
    P = oracle_mod.uniform(8000, 32, 61, lo=-1.0, hi=1.0, dtype=np.float64)
    P = np.round(P * 15.0)                       # exact in E4M3: the FP8 answer is the exact one
    _case("fp8_allknn_f64", m=8000, d=32, k=8)
    D, I = oracle_mod.exact_topk(P[:500], P, 8, self_ids=np.arange(500))
    with knns_mod.Corpus(P, fp8=True) as c:
        dd, ii = c.search(P[:500], 8, precision="fp8", exclude_self=True, self_offset=0)
    assert np.array_equal(ii.astype(np.int64), I.astype(np.int64))
    assert np.array_equal(dd, D)


@pytest.mark.gpu
def test_fp8_needs_fp8_corpus(gpu, knns_mod):
    R = np.zeros((100, 16), np.float32)
    with knns_mod.Corpus(R) as c:
        with pytest.raises(knns_mod.KnnsError):
            c.search(R[:4], 3, precision="fp8")


def test_fp8_plan_uses_fp8_row_bytes(knns_mod, builtin_plans):
    # the FP8 scan's chunk is the FP8 row (d rounded up to 64, 128 or 256 bytes, the widths the
    # kernel is built for); the FP32 rerank rows keep their own padding
    for d, dc, dp in [(1, 64, 32), (24, 64, 32), (64, 64, 64), (65, 128, 96), (96, 128, 96), (128, 128, 128),
                      (129, 256, 256), (130, 256, 256), (200, 256, 256), (256, 256, 256)]:
        p = knns_mod.plan_describe(100_000, d, 1000, 10, "fp8")
        assert (p["dc"], p["dp"], p["qb"]) == (dc, dp, 2), d
    assert knns_mod.plan_describe(100_000, 96, 20, 10, "fp8")["qb"] == 1
