"""Test configuration.

Markers: ``gpu`` marks tests that need a ROCm GPU (they call the HIP kernels through the C ABI).
The CPU tier (``-m "not gpu"``) covers the oracle against closed forms and golden vectors, the host
logic of the Python boundary and the C-ABI library's exports and argument validation.
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (gfx950); runs the HIP kernels")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    return torch.device("cuda", 0)
