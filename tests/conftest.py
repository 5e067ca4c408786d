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
