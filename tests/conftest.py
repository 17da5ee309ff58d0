"""Test configuration: import paths, the `gpu` marker, shared helpers.

CPU tests (``-m "not gpu"``) pin the oracle (oracle/) against the golden
vectors of the reference's tests and independent checks, cover the host
logic and check that libstencils_hip.so exports the ABI.  GPU tests call the
HIP path through the C ABI and compare it with the oracle.
"""
import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "stencils.jl_amd"))
sys.path.insert(0, str(REPO / "oracle"))
sys.path.insert(0, str(REPO / "tests"))

# Two-step launches run by default only on large arrays; tests run them
# on every eligible array so small shapes cover their boundary cases.
os.environ.setdefault("STH_FUSE3D_MIN_CELLS", "0")
os.environ.setdefault("STH_FUSE2D_MIN_CELLS", "0")

GOLDEN = json.loads((REPO / "tests" / "golden" / "reference_goldens.json").read_text())


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libstencils_hip.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.build()
    return orc


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def gpu():
    """The product path; fails (does not skip) when the library or device is missing."""
    import torch
    import stencils_hip as sh
    sh.load_library()
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return sh


def colmajor(a, dtype=None):
    return np.asfortranarray(np.asarray(a, dtype=dtype))
