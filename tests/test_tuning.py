"""Planner tuning table (knn-toolbox_amd/tuning/gfx950.tsv, written by knn-toolbox_amd/tune.py):
parsing, nearest-entry lookup and its effect on the plan. CPU only (knns_plan_describe does no
device work)."""
from pathlib import Path

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
        d, k, n, m, w, ratio, qb = map(int, r[1:8])
        assert d > 0 and 1 <= k <= 1024 and n > 0 and m > 0 and w >= 1 and (ratio == 0 or ratio >= 2)
        assert 0 <= qb <= 4
        assert float(r[8]) > 0


def test_table_drives_slices_and_ratio(knns_mod, tmp_path, restore_table):
    L = knns_mod.native()
    # 10k queries of 64 (QB=2 at BF16): 40 workgroup rows; with 512 resident workgroups the
    # smallest slice count filling >= w waves at >= 95% is 52..64 (w=5) and 39..52 (w=4), ...
    slices = {}
    for w in (3, 5, 8):
        t = tmp_path / f"w{w}.tsv"
        t.write_text(f"# test\nbf16 128 100 10000 1000000 {w} 0 0\n")
        assert L.knns_set_tuning_table(str(t).encode()) == 0
        slices[w] = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["slices_max"]
    assert slices[3] < slices[5] < slices[8]
    # ratio: fixed x2 growth gives more rounds than the default schedule, x8 fewer
    rounds = {}
    for r in (0, 2, 8):
        t = tmp_path / f"r{r}.tsv"
        t.write_text(f"bf16 128 100 10000 1000000 5 {r} 0\n")
        assert L.knns_set_tuning_table(str(t).encode()) == 0
        rounds[r] = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["rounds"]
    assert rounds[2] > rounds[0] > rounds[8]


def test_nearest_entry_and_precision(knns_mod, tmp_path, restore_table):
    L = knns_mod.native()
    t = tmp_path / "t.tsv"
    t.write_text("bf16 128 100 10000 1000000 3 0 0\nbf16 128 100 256 1000000 8 0 0\nfp32 128 100 10000 1000000 8 0 0\n")
    assert L.knns_set_tuning_table(str(t).encode()) == 0
    near_big = knns_mod.plan_describe(1_000_000, 120, 9000, 90, "bf16")["slices_max"]
    t2 = tmp_path / "t2.tsv"
    t2.write_text("bf16 128 100 10000 1000000 3 0 0\n")
    assert L.knns_set_tuning_table(str(t2).encode()) == 0
    assert knns_mod.plan_describe(1_000_000, 120, 9000, 90, "bf16")["slices_max"] == near_big
    # cleared table: built-in defaults (5 waves)
    assert L.knns_set_tuning_table(None) == 0
    t3 = tmp_path / "t3.tsv"
    t3.write_text("bf16 128 100 10000 1000000 5 0 0\n")
    default = knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["slices_max"]
    assert L.knns_set_tuning_table(str(t3).encode()) == 0
    assert knns_mod.plan_describe(1_000_000, 128, 10_000, 100, "bf16")["slices_max"] == default


@pytest.mark.parametrize("text,where", [("bf16 128 100\n", ":1:"), ("# c\nfp99 1 1 1 1 5 0 0\n", ":2: unknown precision"),
                                        ("bf16 128 100 10 100 0 0 0\n", "out of range"), ("bf16 128 100 10 100 5 1 0\n", "out of range"),
                                        ("bf16 128 100 10 100 5 0 7\n", "out of range"), ("bf16 128 100 10 0 5 0 0\n", "out of range")])
def test_malformed_tables_rejected(knns_mod, tmp_path, restore_table, text, where):
    L = knns_mod.native()
    t = tmp_path / "bad.tsv"
    t.write_text(text)
    assert L.knns_set_tuning_table(str(t).encode()) == knns_mod.KNNS_ERR_INVALID
    assert where in L.knns_last_error().decode()
    assert L.knns_set_tuning_table(b"/nonexistent/table.tsv") == knns_mod.KNNS_ERR_INVALID
