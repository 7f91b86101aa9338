import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU run (minutes); deselect with -m 'not slow'")


@pytest.fixture(scope="session")
def nu_healthy():
    """Lattice viscosity of the reference's healthy operating point: blood, dx 0.1 mm, dt 5 us (tau 0.50495)."""
    return (3.5e-3 / 1060.0) * 5e-6 / (1e-4) ** 2
