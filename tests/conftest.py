import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running parity case")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The suite library, initialised on device 0 (GPU tests only)."""
    if not gpu_available():
        pytest.skip("no GPU")
    from tealeaf_amd.abi import lib
    L = lib()
    L.tlhip_init(0)
    return L


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import load_oracle
    return load_oracle()
