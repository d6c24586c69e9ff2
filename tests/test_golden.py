"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py with numpy alone).

Integer-valued inputs make every squared distance exact, so the expected answer is unambiguous:
ascending distance, ties to the lower index (demo/knn.m:5 stable sort), +inf / -1 padding past
the corpus. CPU tests pin the fixtures themselves (pure-Python integer arithmetic) and the oracle
against them; GPU tests pin the drop-in libraries and the native precisions."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    out = {key: z[key] for key in z.files}
    out["k"] = int(out["k"])
    if "self_offset" in out:
        out["self_offset"] = int(out["self_offset"])
    return out


def test_fixture_set_is_complete():
    assert {"int_uniform_f32", "int_ties_f32", "int_small_corpus_f32", "int_uniform_f64",
            "int_allknn_f64", "int_bf16_exact_f32"} <= set(CASES)


@pytest.mark.parametrize("name", CASES)
def test_fixture_matches_integer_recomputation(name):
    """Recompute the first queries with Python integers (no floating point at all)."""
    c = load(name)
    R = c["corpus"].astype(np.int64).tolist()
    m, k = len(R), c["k"]
    so = c.get("self_offset")
    for i, q in enumerate(c["queries"].astype(np.int64).tolist()[:3]):
        d2 = [(sum((a - b) * (a - b) for a, b in zip(q, r)), j) for j, r in enumerate(R)
              if so is None or j != so + i]
        d2.sort()
        want = d2[:k]
        got_idx = c["idx"][i].tolist()
        got_d = c["dist"][i].tolist()
        for t in range(k):
            if t < len(want):
                assert (got_d[t], got_idx[t]) == (float(want[t][0]), want[t][1])
            else:
                assert got_idx[t] == -1 and got_d[t] == float("inf")


@pytest.mark.parametrize("name", CASES)
def test_oracle_exact_topk_matches_fixture(oracle_mod, name):
    c = load(name)
    Q, R, k = c["queries"], c["corpus"], c["k"]
    self_ids = None
    if "self_offset" in c:
        self_ids = np.arange(Q.shape[0], dtype=np.int64) + c["self_offset"]
    if k > R.shape[0]:
        kk = R.shape[0]
        D, I = oracle_mod.exact_topk(Q, R, kk, self_ids)
        assert np.array_equal(I, c["idx"][:, :kk]) and np.array_equal(D, c["dist"][:, :kk])
        return
    D, I = oracle_mod.exact_topk(Q, R, k, self_ids)
    assert np.array_equal(I, c["idx"])
    assert np.array_equal(D, c["dist"])


@pytest.mark.parametrize("name", [n for n in CASES if n not in ("int_allknn_f64", "int_small_corpus_f32")])
def test_oracle_knnm_matches_fixture(oracle_mod, name):
    """knn.m transcription (FP32/FP64 key ||y||^2 - 2 x.y, stable sort): exact on these inputs."""
    c = load(name)
    D, I = oracle_mod.knnm(c["queries"], c["corpus"], c["k"])
    assert np.array_equal(I, c["idx"])
    assert np.array_equal(D.astype(np.float64), c["dist"])


# ---------------------------------------------------------------------------------------------
# GPU: the drop-in libraries and the native precisions against the same fixtures
# ---------------------------------------------------------------------------------------------
def _print_case(name, **kw):
    print(f"\n[case] golden {name} " + " ".join(f"{a}={b}" for a, b in kw.items()), flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in CASES if n != "int_allknn_f64"])
@pytest.mark.parametrize("entry", ["heap", "bitonic"])
def test_dropin_matches_fixture(gpu, knns_mod, name, entry):
    c = load(name)
    double = c["corpus"].dtype == np.float64
    _print_case(name, entry=entry, double=double)
    fn = knns_mod.gpuknns_heap if entry == "heap" else knns_mod.gpuknns_bitonic
    vals, inds, t = fn(c["queries"], c["corpus"], c["k"], double=double)
    assert np.array_equal(inds.astype(np.int64), c["idx"].astype(np.int64))
    kt = np.float64 if double else np.float32
    assert np.array_equal(vals, c["dist"].astype(kt))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["int_uniform_f32", "int_ties_f32", "int_small_corpus_f32", "int_bf16_exact_f32"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_native_precisions_match_fixture(gpu, knns_mod, name, precision):
    """Values <= 256 are exact in BF16, so the BF16 shortlist is exact here as well."""
    c = load(name)
    _print_case(name, precision=precision)
    with knns_mod.Corpus(c["corpus"], bf16=precision == "bf16") as corpus:
        D, I = corpus.search(c["queries"], c["k"], precision=precision)
    assert np.array_equal(I, c["idx"])
    assert np.array_equal(D, c["dist"].astype(np.float32))


@pytest.mark.gpu
def test_fp64_allknn_matches_fixture(gpu, knns_mod):
    c = load("int_allknn_f64")
    _print_case("int_allknn_f64", precision="fp64")
    with knns_mod.Corpus(c["corpus"]) as corpus:
        D, I = corpus.search(c["queries"], c["k"], precision="fp64", exclude_self=True,
                             self_offset=c["self_offset"])
    assert np.array_equal(I, c["idx"])
    assert np.array_equal(D, c["dist"])
