"""The knnTest driver (src/knnTest.cu:65-110): raw .bin files in, KNNdist.bin / KNNidx.bin out.

CPU tests cover argument and file checks (the program exits before any device work); the GPU
tests run a search through the files and compare with the FP64 oracle like the drop-in tests
(indices identical, distances equal to the oracle's FP64 value rounded to knntype)."""
import subprocess

import numpy as np
import pytest


def _run(exe, args, cwd):
    return subprocess.run([str(exe), *map(str, args)], cwd=str(cwd), capture_output=True, text=True,
                          timeout=120)


@pytest.fixture(scope="module")
def exe(knns_mod):
    p = knns_mod.knntest_path(False)
    if not p.exists():
        pytest.fail(f"{p} missing: run python knn-toolbox_amd/build.py")
    return p


def test_usage_without_arguments(exe, tmp_path):
    r = _run(exe, [], tmp_path)
    assert r.returncode == 2
    assert "usage" in r.stderr


@pytest.mark.parametrize("args", [["a", "b", "x", "2", "3", "1"], ["a", "b", "10", "0", "3", "1"],
                                  ["a", "b", "10", "2", "3", "2000"], ["a", "b", "10", "2", "-1", "1"]])
def test_rejects_bad_sizes(exe, tmp_path, args):
    r = _run(exe, args, tmp_path)
    assert r.returncode == 2
    assert "must be integers" in r.stderr


def test_missing_and_short_files(exe, knns_mod, tmp_path):
    r = _run(exe, ["nope.bin", "nope2.bin", 10, 2, 3, 1], tmp_path)
    assert r.returncode == 2 and "cannot open input file" in r.stderr
    knns_mod.export_bin(tmp_path / "c.bin", np.zeros((9, 3)))       # 27 values, 30 expected
    knns_mod.export_bin(tmp_path / "q.bin", np.zeros((2, 3)))
    r = _run(exe, ["c.bin", "q.bin", 10, 2, 3, 1], tmp_path)
    assert r.returncode == 2 and "expected 30" in r.stderr
    assert not (tmp_path / "KNNdist.bin").exists()


def test_bin_helpers_round_trip(knns_mod, tmp_path):
    # importData.m reads fread(fid, [k Q]): k leading, i.e. Q rows of k values
    Q, k = 3, 4
    d = np.arange(Q * k, dtype=np.float32).reshape(Q, k)
    d.tofile(str(tmp_path / "KNNdist.bin"))
    (d + 100).tofile(str(tmp_path / "KNNidx.bin"))
    D, I = knns_mod.import_results(Q, k, directory=tmp_path)
    assert np.array_equal(D, d) and np.array_equal(I, d + 100)
    with pytest.raises(knns_mod.KnnsError):
        knns_mod.import_results(Q, k + 1, directory=tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("double", [False, True])
def test_knntest_matches_oracle(gpu, knns_mod, oracle_mod, tmp_path, double):
    m, n, d, k = 3000, 50, 40, 8
    print(f"\n[case] knnTest double={double} m={m} n={n} d={d} k={k}", flush=True)
    kt = np.float64 if double else np.float32
    R = oracle_mod.uniform(m, d, seed=21, dtype=kt)
    Q = oracle_mod.uniform(n, d, seed=22, dtype=kt)
    knns_mod.export_bin(tmp_path / "TestData.bin", R, double)
    knns_mod.export_bin(tmp_path / "QueryData.bin", Q, double)
    r = _run(knns_mod.knntest_path(double), ["TestData.bin", "QueryData.bin", m, n, d, k], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Time Elapsed:" in r.stdout and "Saving data to file: KNNidx.bin" in r.stdout
    D, I = knns_mod.import_results(n, k, double, tmp_path)
    Do, Io = oracle_mod.exact_topk(Q, R, k)
    assert np.array_equal(I.astype(np.int64), Io.astype(np.int64))
    assert np.array_equal(D, Do.astype(kt))
