import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

EMU_LIB = os.path.join(ROOT, "tests", "cpuemu", "libhfx_cpuemu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libhfx.so on the device)")
    config.addinivalue_line("markers", "slow: long CPU test")


def core_guess(s, h, nocc):
    """Core-Hamiltonian guess density, as direct_hartreefock.f90:24-29,66-72 builds it."""
    w, U = np.linalg.eigh(s)
    X = U @ np.diag(w ** -0.5) @ U.T
    e, C = np.linalg.eigh(X @ h @ X.T)
    C = X @ C
    return 2.0 * C[:, :nocc] @ C[:, :nocc].T


@pytest.fixture
def emu_fock(monkeypatch):
    """hfcode_amd.fock bound to the CPU emulation build of libhfx (tests only)."""
    from hfcode_amd import fock
    if not os.path.exists(EMU_LIB):
        from tests.cpuemu import build_cpuemu
        build_cpuemu.build()
    monkeypatch.setattr(fock, "_lib", fock.configure(ctypes.CDLL(EMU_LIB)))
    return fock
