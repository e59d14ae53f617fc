"""hfx_ctx_onee_noalloc: S, T, V with every device buffer in the library's
fixed scratch array (no allocation in the call), against hfx_ctx_onee on the
same context: two calls in a row, the second on a smaller molecule, and a
molecule beyond the scratch size refused with an error."""
import ctypes

import numpy as np
import pytest

from hfcode_amd import fock, molecule as M

pytestmark = pytest.mark.gpu


def _noalloc(fb, mol):
    L = fock.lib()
    nb = fb.nb
    ch = np.ascontiguousarray(mol.znuc, dtype=np.float64)
    xy = np.ascontiguousarray(mol.coords_bohr, dtype=np.float64).reshape(-1, 3)
    S, T, V = (np.full((nb, nb), np.nan) for _ in range(3))
    rc = L.hfx_ctx_onee_noalloc(fb._h, len(ch), fock._dp(ch), fock._dp(xy), fock._dp(S), fock._dp(T), fock._dp(V))
    return rc, S, T, V


def test_noalloc_onee_matches_allocating_entry_point():
    cases = [(M.c60(), "6-31g(d)"), (M.benzene(), "6-31g(d)"), (M.water(), "sto-3g")]   # large, then smaller
    for mol, basis in cases:
        fb = fock.FockBuilder(M.ShellTable(mol, M.load_basis(basis)))
        try:
            rc, S, T, V = _noalloc(fb, mol)
            assert rc == 0, (mol.title, rc)
            S0, T0, V0 = fb.onee(mol.znuc.astype(np.float64), mol.coords_bohr)
            for a, b in ((S, S0), (T, T0), (V, V0)):
                assert np.array_equal(a, b), mol.title
            print("%s: nb %d, no-alloc S/T/V equal to hfx_ctx_onee" % (mol.title, fb.nb), flush=True)
        finally:
            fb.close()


def test_noalloc_onee_refuses_oversized_molecule():
    mol = M.chain(120, "alkane")          # 2164 basis functions in 6-31G(d): S, T, V exceed the scratch
    fb = fock.FockBuilder(M.ShellTable(mol, M.load_basis("6-31g(d)")))
    try:
        rc, S, T, V = _noalloc(fb, mol)
        assert rc == -2                   # HFX_E_UNSUPPORTED
        assert np.isnan(S).all()          # nothing written
    finally:
        fb.close()
