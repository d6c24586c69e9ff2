"""Planner tuning table (knn-toolbox_amd/tuning/gfx950.tsv, written by knn-toolbox_amd/tune.py):
parsing, nearest-entry lookup and its effect on the plan. CPU only (knns_plan_describe does no
device work)."""
from pathlib import Path

import numpy as np
import pytest

TABLE = Path(__file__).resolve().parent.parent / "knn-toolbox_amd" / "tuning" / "gfx950.tsv"


@pytest.fixture()
def restore_table(knns_mod):
    yield
    assert knns_mod.native().knns_set_tuning_table(str(TABLE).encode()) == 0


def test_committed_table_parses(knns_mod, restore_table):
    L = knns_mod.native()
    assert L.knns_set_tuning_table(str(TABLE).encode()) == 0
    rows = [l.split() for l in TABLE.read_text().splitlines() if l.strip() and not l.startswith("#")]
    assert len(rows) >= 60
    assert {r[0] for r in rows} == {"bf16", "fp32", "fp8"}
    # every entry: d, k, n and the two knobs are integers in range, the time is positive
    for r in rows:
        d, k, n, m, w, ratio, qb, extra = map(int, r[1:9])
        assert d > 0 and 1 <= k <= 1024 and n > 0 and m > 0 and w >= 1 and (ratio == 0 or ratio >= 2)
        assert 0 <= qb <= 4 and -1024 <= extra <= 1024
        assert float(r[9]) > 0


def test_table_drives_slices_and_ratio(knns_mod, tmp_path, restore_table):
    L = knns_mod.native()
    # 10k queries of 64 (QB=2 at BF16): 40 workgroup rows; with 512 resident workgroups the
    # smallest slice count filling >= w waves at >= 95% is 52..64 (w=5) and 39..52 (w=4), ...
    slices = {}
    for w in (3, 5, 8):
        t = tmp_path / f"w{w}.tsv"
        t.write_text(f"# test\nbf16 128 100 10000 1000000 {w} 0 0 0\n")
        assert L.knns_set_tuning_table(str(t).encode()) == 0
        slices[w] = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["slices_max"]
    assert slices[3] < slices[5] < slices[8]
    # ratio: fixed x2 growth gives more rounds than the default schedule, x16 fewer
    rounds = {}
    for r in (0, 2, 16):
        t = tmp_path / f"r{r}.tsv"
        t.write_text(f"bf16 128 100 10000 1000000 5 {r} 0 0\n")
        assert L.knns_set_tuning_table(str(t).encode()) == 0
        rounds[r] = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["rounds"]
    assert rounds[2] > rounds[0] > rounds[16]


def test_nearest_entry_and_precision(knns_mod, tmp_path, restore_table):
    L = knns_mod.native()
    t = tmp_path / "t.tsv"
    t.write_text("bf16 128 100 10000 1000000 3 0 0 0\nbf16 128 100 256 1000000 8 0 0 0\nfp32 128 100 10000 1000000 8 0 0 0\n")
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    near_big = knns_mod.plan_describe(1_000_000, 120, 9000, 90, "bf16")["slices_max"]
    t2 = tmp_path / "t2.tsv"
    t2.write_text("bf16 128 100 10000 1000000 3 0 0 0\n")
    assert L.knns_set_tuning_table(str(t2).encode()) == 0
    assert knns_mod.plan_describe(1_000_000, 120, 9000, 90, "bf16")["slices_max"] == near_big
    # cleared table: built-in defaults (5 waves)
    assert L.knns_set_tuning_table(None) == 0
    t3 = tmp_path / "t3.tsv"
    t3.write_text("bf16 128 100 10000 1000000 5 0 0 0\n")
    default = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["slices_max"]
    assert L.knns_set_tuning_table(str(t3).encode()) == 0
    assert knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["slices_max"] == default


def test_fp64_searches_use_fp32_entries_without_their_own(knns_mod, tmp_path, restore_table):
    """FP64 searches run the FP32 scan and shortlist rule: with no fp64 entry in the table they
    take the nearest fp32 entry; an fp64 entry, when present, wins."""
    L = knns_mod.native()
    t = tmp_path / "f.tsv"
    t.write_text("fp32 32 10 10000 1000000 5 0 2 -16\n")
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    p = knns_mod.plan_describe(200_000, 32, 200_000, 8, "fp64")
    assert p["qb"] == 2 and p["shortlist"] == 8          # k + 16 - 16
    t.write_text("fp32 32 10 10000 1000000 5 0 2 -16\nfp64 32 10 10000 1000000 5 0 1 0\n")
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    p = knns_mod.plan_describe(200_000, 32, 200_000, 8, "fp64")
    assert p["qb"] == 1 and p["shortlist"] == 24
    assert L.knns_set_tuning_table(None) == 0
    assert knns_mod.plan_describe(200_000, 32, 200_000, 8, "fp64")["qb"] == 1


@pytest.mark.parametrize("text,where", [("bf16 128 100\n", ":1:"), ("# c\nfp99 1 1 1 1 5 0 0 0\n", ":2: unknown precision"),
                                        ("bf16 128 100 10 100 0 0 0 0\n", "out of range"), ("bf16 128 100 10 100 5 1 0 0\n", "out of range"),
                                        ("bf16 128 100 10 100 5 0 7 0\n", "out of range"), ("bf16 128 100 10 0 5 0 0 0\n", "out of range"),
                                        ("bf16 128 100 10 100 5 0 0 2000\n", "out of range")])
def test_malformed_tables_rejected(knns_mod, tmp_path, restore_table, text, where):
    L = knns_mod.native()
    t = tmp_path / "bad.tsv"
    t.write_text(text)
    assert L.knns_set_tuning_table(str(t).encode()) == knns_mod.KNNS_ERR_INVALID
    assert where in L.knns_last_error().decode()
    assert L.knns_set_tuning_table(b"/nonexistent/table.tsv") == knns_mod.KNNS_ERR_INVALID


def test_table_drives_tile_shape_and_shortlist(knns_mod, tmp_path, restore_table):
    """The tile shape (query blocks per scan wave) and the shortlist come from the table; a shape
    the scan is not built for falls back to the built-in rule; corpus size is part of the key."""
    L = knns_mod.native()
    got = {}
    for qb in (1, 2, 3, 4):
        t = tmp_path / f"qb{qb}.tsv"
        t.write_text(f"bf16 128 100 10000 1000000 5 0 {qb} 0\n")
        assert L.knns_set_tuning_table(str(t).encode()) == 0
        got[qb] = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")
    assert [got[q]["qb"] for q in (1, 2, 3, 4)] == [1, 2, 3, 4]
    # fewer queries per workgroup row at smaller QB: more query groups
    assert got[1]["groups"] > got[2]["groups"] > got[4]["groups"]
    t = tmp_path / "fp32.tsv"
    t.write_text("fp32 128 100 10000 1000000 5 0 4 0\n")          # FP32 has no 4-block scan
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    assert knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "fp32")["qb"] == 1
    # shortlist offset
    base = None
    for extra in (0, 40, -60):
        t = tmp_path / f"sl{extra}.tsv"
        t.write_text(f"bf16 128 100 10000 1000000 5 0 2 {extra}\n")
        assert L.knns_set_tuning_table(str(t).encode()) == 0
        sl = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["shortlist"]
        base = sl if extra == 0 else base
        assert sl == base + extra
    # corpus size as a key: two entries differing only in m
    t = tmp_path / "m.tsv"
    t.write_text("bf16 128 100 10000 1000000 5 0 1 0\nbf16 128 100 10000 10000000 5 0 2 0\n")
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    assert knns_mod.plan_describe(1_200_000, 128, 10_000, 100, "bf16")["qb"] == 1
    assert knns_mod.plan_describe(9_000_000, 128, 10_000, 100, "bf16")["qb"] == 2


@pytest.mark.gpu
def test_tile_shapes_give_identical_results(gpu, knns_mod, oracle_mod, tmp_path, restore_table):
    """Every tile shape the table can select returns the same answer (the oracle's)."""
    L = knns_mod.native()
    R = oracle_mod.mixture(200000, 128, 128, 0.1, center_seed=61, point_seed=62)
    Q = oracle_mod.mixture(700, 128, 128, 0.1, center_seed=61, point_seed=63)
    Do, Io = oracle_mod.exact_topk(Q[:64], R, 50)
    with knns_mod.Corpus(R, bf16=True) as c:
        ref = None
        for qb in (1, 2, 3, 4):
            t = tmp_path / f"g{qb}.tsv"
            t.write_text(f"bf16 128 50 700 200000 5 0 {qb} 0\n")
            assert L.knns_set_tuning_table(str(t).encode()) == 0
            print(f"\n[case] tile qb={qb}", flush=True)
            D, I = c.search(Q, 50, precision="bf16")
            assert knns_mod.last_stats()["qb"] == qb
            assert np.array_equal(I[:64], Io) and np.array_equal(D[:64], Do.astype(np.float32))
            if ref is None:
                ref = (D, I)
            assert np.array_equal(I, ref[1]) and np.array_equal(D, ref[0])


@pytest.mark.gpu
@pytest.mark.parametrize("prec,d", [("fp8", 96), ("fp8", 64), ("bf16", 64)])
def test_wide_tiles_give_identical_results(gpu, knns_mod, oracle_mod, tmp_path, restore_table, prec, d):
    """Three and four query blocks per wave (the pipelined scan of FP8 rows up to 128 bytes and of
    64-dim BF16 rows; four blocks with one accumulator set) compute the same keys as one and two:
    identical answers, and for BF16 the oracle's."""
    L = knns_mod.native()
    R = oracle_mod.mixture(150000, d, 128, 0.1, center_seed=81, point_seed=82)
    Q = oracle_mod.mixture(900, d, 128, 0.1, center_seed=81, point_seed=83)
    k = 20
    Do, Io = oracle_mod.exact_topk(Q[:48], R, k)
    with knns_mod.Corpus(R, bf16=True, fp8=prec == "fp8") as c:
        ref = None
        for qb in (1, 2, 3, 4):
            t = tmp_path / f"t{qb}.tsv"
            t.write_text(f"{prec} {d} {k} 900 150000 5 0 {qb} 0\n")
            assert L.knns_set_tuning_table(str(t).encode()) == 0
            print(f"\n[case] {prec} d={d} tile qb={qb}", flush=True)
            D, I = c.search(Q, k, precision=prec)
            assert knns_mod.last_stats()["qb"] == qb
            if prec == "bf16":
                assert np.array_equal(I[:48], Io) and np.array_equal(D[:48], Do.astype(np.float32))
            else:
                assert oracle_mod.recall_at_k(I[:48], Io) > 0.9
            if ref is None:
                ref = (D, I)
            assert np.array_equal(I, ref[1]) and np.array_equal(D, ref[0])


@pytest.mark.gpu
def test_server_shortlist_follows_each_k(gpu, knns_mod, oracle_mod, tmp_path, restore_table):
    """A server's shortlist for a request of k is the one a plain search of that k would use: the
    table's offset for max_k (here -32 at k = 100) is not applied to k = 1 requests, whose own
    rule (k + 16, with the k = 10 entry's offset 0) certifies every query in one pass."""
    L = knns_mod.native()
    t = tmp_path / "srv.tsv"
    t.write_text("bf16 128 100 4096 200000 5 0 2 -32\nbf16 128 10 4096 200000 5 0 2 0\n")
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    R = oracle_mod.mixture(200000, 128, 128, 0.1, center_seed=71, point_seed=72)
    Q = oracle_mod.mixture(3000, 128, 128, 0.1, center_seed=71, point_seed=73)
    Do, Io = oracle_mod.exact_topk(Q[:32], R, 1)
    with knns_mod.Corpus(R, bf16=True) as c:
        print("\n[case] server k=1 with max_k=100", flush=True)
        with knns_mod.Server(c, precision="bf16", max_k=100, max_batch=4096) as srv:
            srv.stats(reset=True)
            tickets = [srv.submit(Q[s:s + 500], 1) for s in range(0, 3000, 500)]
            got = [srv.result(tk) for tk in tickets]
            st = srv.stats()
        assert st["researched"] == 0
        D, I = np.concatenate([g[0] for g in got]), np.concatenate([g[1] for g in got])
        assert np.array_equal(I[:32], Io) and np.array_equal(D[:32], Do.astype(np.float32))
        Da, Ia = c.search(Q, 1, precision="bf16")
        assert knns_mod.last_stats()["uncertified"] == 0
        assert np.array_equal(Ia, I) and np.array_equal(Da, D)
