"""Shared pytest configuration.

Markers:
  gpu  -- needs a ROCm GPU and the built libjfm.so; parity against the oracle.
Everything unmarked runs on the CPU (oracle pins, host topology, ABI surface).
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libjfm.so")


def gpu_available() -> bool:
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test collected on a machine without a GPU (run with -m 'not gpu')")
