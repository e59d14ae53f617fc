"""ORACLE-SIDE TEST INFRASTRUCTURE -- not part of the product.

ctypes access to the C restatement oracle/_ref/libhfx_oracle.so
(oracle/hfx_oracle.c): Boys function as the reference evaluates it
(boys.f90), McMurchie--Davidson integral blocks with the reference's
primitive-pair screening, and the four Fock builds with the reference's
screening and digestion rules.  Inputs use the hfx_mol_desc layout of
include/hfx.h.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libhfx_oracle.so")
KINDS = {"non_diag": 0, "j_only": 1, "diag": 2, "initial_diag": 3}
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError("%s not built (make -C oracle oracle)" % LIB_PATH)
        from hfcode_amd.fock import MolDesc
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.oracle_boys.restype = None
        L.oracle_fock.argtypes = [ctypes.POINTER(MolDesc), dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.oracle_eri_block.argtypes = [ctypes.POINTER(MolDesc)] + [ctypes.c_int] * 5 + [dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def boys(n, x):
    out = np.zeros(n + 1)
    lib().oracle_boys(int(n), float(x), _dp(out))
    return out


def fock(desc, P, F, thr, kind, ntot_pairs=0):
    """F + G for `kind` in KINDS; returns (F_new, schw_out or None)."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    F = np.array(F, dtype=np.float64, copy=True, order="C")
    so = np.zeros(max(ntot_pairs, 1))
    rc = lib().oracle_fock(ctypes.byref(desc), _dp(P), _dp(F), float(thr), KINDS[kind], _dp(so))
    if rc:
        raise RuntimeError("oracle_fock failed")
    return F, (so[:ntot_pairs] if kind == "initial_diag" else None)


def eri_block(desc, s1, s2, s3, s4, dims, skip_negab=True):
    out = np.zeros(625)
    lib().oracle_eri_block(ctypes.byref(desc), s1, s2, s3, s4, int(skip_negab), _dp(out))
    return out[:int(np.prod(dims))].reshape(dims)
