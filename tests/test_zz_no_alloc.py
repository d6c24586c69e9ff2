"""knns_search_noalloc (no device or pinned allocation; workspace from the library's fixed device
scratch array) against knns_search with device pointers, through the C ABI: identical indices and
distances on two calls in a row (the second smaller, so it reuses the array's stale contents), for
the FP32 and BF16 forms, and the size limit reported as an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"\n[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.fixture(scope="module")
def corpus(gpu, knns_mod, oracle_mod):
    R = oracle_mod.mixture(60000, 128, 256, 0.1, center_seed=81, point_seed=82)
    c = knns_mod.Corpus(R, bf16=True)
    yield R, c
    c.close()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_noalloc_equals_search_on_two_calls(knns_mod, oracle_mod, corpus, prec):
    torch = pytest.importorskip("torch")
    R, c = corpus
    Q = oracle_mod.mixture(3000, 128, 256, 0.1, center_seed=81, point_seed=83)
    for n, k in ((3000, 50), (700, 10)):            # the second call is smaller
        _case("noalloc", prec=prec, n=n, k=k)
        q = _dev(torch, Q[:n])
        d_ref = torch.empty((n, k), dtype=torch.float32, device="cuda:0")
        i_ref = torch.empty((n, k), dtype=torch.int32, device="cuda:0")
        d_na = torch.full((n, k), -1.0, dtype=torch.float32, device="cuda:0")
        i_na = torch.full((n, k), -7, dtype=torch.int32, device="cuda:0")
        c.search_device(q, n, k, d_ref, i_ref, precision=prec)
        c.search_noalloc(q, n, k, d_na, i_na, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(i_ref, i_na) and torch.equal(d_ref, d_na)
        # and the answer is the oracle's on a sample
        s = np.arange(0, n, max(1, n // 16))
        Do, Io = oracle_mod.exact_topk(Q[s], R, k)
        assert np.array_equal(i_na.cpu().numpy()[s], Io) and np.array_equal(d_na.cpu().numpy()[s], Do.astype(np.float32))


def test_noalloc_limits_and_arguments(knns_mod, corpus):
    torch = pytest.importorskip("torch")
    R, c = corpus
    n, k = 100000, 100                               # one slice's sub-pools alone exceed the array
    _case("noalloc_limit", n=n, k=k)
    q = torch.zeros((n, 128), dtype=torch.float32, device="cuda:0")
    d = torch.empty((n, k), dtype=torch.float32, device="cuda:0")
    i = torch.empty((n, k), dtype=torch.int32, device="cuda:0")
    with pytest.raises(knns_mod.KnnsError) as e:
        c.search_noalloc(q, n, k, d, i, precision="bf16")
    assert e.value.status == knns_mod.KNNS_ERR_NOMEM and "scratch" in str(e.value)
    # host pointers are refused before any device work
    p = c._params("fp32", 10, "auto", 0, False, 0, 0, 0, 0)
    qh = np.zeros((4, 128), dtype=np.float32)
    dh = np.zeros((4, 10), dtype=np.float32)
    ih = np.zeros((4, 10), dtype=np.int32)
    st = knns_mod.native().knns_search_noalloc(c._h, qh.ctypes.data, knns_mod.KNNS_F32, 4,
                                               knns_mod.ctypes.byref(p), dh.ctypes.data, ih.ctypes.data)
    assert st == knns_mod.KNNS_ERR_INVALID
