import sys
from pathlib import Path

import pytest

# torch ships its own HIP runtime beside the system one the library links; the two coexist in
# one process when torch is imported before the library is first loaded (the order of bench.py),
# so tests that hand torch device tensors to the library see the same device either way
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "oracle", ROOT / "knn-toolbox_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on a machine that has one)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def knns_mod():
    import knns

    return knns


@pytest.fixture(scope="session")
def gpu(knns_mod):
    n = knns_mod.device_count()
    if n <= 0:
        pytest.fail("no GPU visible to the HIP runtime (GPU tests need an MI355X)")
    return n


@pytest.fixture
def builtin_plans(knns_mod):
    """The planner's built-in rules (no tuning table); the package's table is restored afterwards."""
    L = knns_mod.native()
    assert L.knns_set_tuning_table(None) == 0
    yield
    table = ROOT / "knn-toolbox_amd" / "tuning" / "gfx950.tsv"
    assert L.knns_set_tuning_table(str(table).encode()) == 0
