import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libgcge_ref.so")
REF_SRC = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The HIP library, loaded; fails (never skips) when it is missing, so a
    GPU run cannot pass without the native code."""
    from gcge_amd import abi
    lib = abi.load()
    return lib


@pytest.fixture(scope="session")
def reflib():
    """The compiled reference (oracle/_ref), or skip when it was not built."""
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (needs the reference sources and MKL)")
    from ref_binding import RefLib
    try:
        return RefLib(REF_LIB)
    except OSError as e:
        pytest.skip(f"oracle/_ref does not load here: {e}")
