import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU or GPU runs")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def deck_spec():
    from epi_amd import deck
    return deck.load_npz(os.path.join(GOLDEN, "deck_2layer_nfix16.npz"))


# GPU tests run from the most basic to the most composite: single-tissue
# parity first, then the windows and their adjoints, then batches.  Within a
# module the file order is kept.
_GPU_ORDER = ("test_gpu_parity", "test_gpu_rd_adjoint", "test_gpu_mech_adjoint", "test_gpu_batched",
              "test_gpu_configs", "test_zz_no_alloc")


def pytest_collection_modifyitems(config, items):
    def rank(item):
        mod = item.module.__name__.rsplit(".", 1)[-1] if item.module else ""
        return _GPU_ORDER.index(mod) if mod in _GPU_ORDER else -1
    items.sort(key=rank)   # stable: non-GPU modules keep their places first
