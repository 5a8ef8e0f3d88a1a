"""Test configuration: the `gpu` marker and shared helpers.

CPU tests (-m "not gpu") cover the oracle against the golden fixtures and the
compiled reference, the host-side formats, and that librcgrid.so loads and
exports its C ABI.  GPU tests (-m gpu) run the HIP library on the device and
compare with the oracle.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def hip_device():
    """Fails (does not skip) when no HIP device is visible: GPU tests must
    run the HIP library."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return torch.device("cuda:0")
