import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def golden_values():
    import json
    with open(os.path.join(GOLDEN, "golden_values.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import binding
    binding.build()
    return binding.lib()


@pytest.fixture(scope="session")
def engine_mod():
    """The engine's Python binding with the HIP library loaded (no fallback)."""
    from graspa_amd import engine
    engine.lib()
    return engine


def load_golden_system(name):
    from graspa_amd import deck
    return deck.load_system(os.path.join(GOLDEN, name + ".npz"))
