"""Shared fixtures: the reference's invented test networks (tests/golden/data), the gpu marker and device."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

DATA = Path(__file__).resolve().parent / "golden" / "data"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long CPU cases")


@pytest.fixture
def read_network():
    """Read one of the reference's invented networks (tests/data of the reference, copied as fixtures)."""
    from maxichem_amd.kinvulcan import VulcanNetwork

    def read(stem: str = "synthetic"):
        return VulcanNetwork(DATA / f"{stem}_network.txt", DATA / "NASA9")

    return read


@pytest.fixture(scope="session")
def gpu():
    """The HIP device; skips where there is none."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("this test needs a GPU")
    return torch.device("cuda")
