"""Shared fixtures. `gpu` tests need an MI355X and the built library."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libnosh_amd.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def has_gpu():
    return gpu_available()


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def phase_align(psi, ref):
    """psi * exp(-i theta*) with theta* = arg(<ref, psi>) (global gauge)."""
    z = psi[0::2] + 1j * psi[1::2]
    zr = ref[0::2] + 1j * ref[1::2]
    th = np.angle(np.vdot(zr, z))
    w = z * np.exp(-1j * th)
    out = np.empty_like(psi)
    out[0::2] = w.real
    out[1::2] = w.imag
    return out


def pytest_collection_modifyitems(config, items):
    """Without a GPU, `gpu` tests are skipped (with the reason) instead of erroring."""
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible (torch.cuda.is_available() is False)")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
