import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

EMU_LIB = os.path.join(ROOT, "tests", "cpuemu", "libhfx_cpuemu.so")
LANES_LIB = os.path.join(ROOT, "tests", "cpuemu", "libhfx_lanes.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libhfx.so on the device)")
    config.addinivalue_line("markers", "slow: long CPU test")


from hfcode_amd.scf import core_guess  # noqa: E402,F401


@pytest.fixture
def emu_fock(monkeypatch):
    """hfcode_amd.fock bound to the CPU emulation build of libhfx (tests only)."""
    from hfcode_amd import fock
    if not os.path.exists(EMU_LIB):
        from tests.cpuemu import build_cpuemu
        build_cpuemu.build()
    monkeypatch.setattr(fock, "_lib", fock.configure(ctypes.CDLL(EMU_LIB)))
    return fock


@pytest.fixture
def emu_lanes(monkeypatch):
    """hfcode_amd.fock bound to the 64-lane CPU emulation build of libhfx
    (every workgroup as its blockDim.x lanes in 64-lane waves; tests only)."""
    from hfcode_amd import fock
    if not os.path.exists(LANES_LIB):
        from tests.cpuemu import build_cpuemu
        build_cpuemu.build(lanes=True)
    monkeypatch.setattr(fock, "_lib", fock.configure(ctypes.CDLL(LANES_LIB)))
    return fock
