"""bench.py's command-line contract and the synthetic workloads it draws from (CPU only)."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_workloads_are_seeded_and_shared_with_the_oracle(oracle_mod):
    sys.path.insert(0, str(ROOT / "knn-toolbox_amd"))
    import workloads

    # the oracle's generators are the benchmark's (one definition)
    assert oracle_mod.mixture is workloads.mixture and oracle_mod.workload_c2 is workloads.workload_c2
    a = workloads.mixture(1000, 16, 32, 0.1, center_seed=2, point_seed=1002)
    b = workloads.mixture(1000, 16, 32, 0.1, center_seed=2, point_seed=1002)
    assert np.array_equal(a, b) and a.dtype == np.float32
    # the C2 generator at a reduced size is the C2 mixture definition (centre seed 2, point seed 1002)
    R, Q = workloads.workload_c2(m=2000, n=50)
    assert R.shape == (2000, 128) and Q.shape == (50, 128)
    assert np.array_equal(R, workloads.mixture(2000, 128, 1024, 0.1, center_seed=2, point_seed=1002))
    u = workloads.uniform(7, 5, seed=6, lo=-1.0, hi=1.0, dtype=np.float64)
    assert u.dtype == np.float64 and u.min() >= -1.0 and u.max() < 1.0
    s = workloads.log_uniform_sizes(64, 64, 16384, seed=7, total=200_000)
    assert int(s.sum()) == 200_000 and s.min() >= 1
