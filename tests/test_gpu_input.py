"""A job set up only from the reference's input and basis texts (native
readers) reproduces the reference's Fock builds on the GPU."""
import numpy as np
import pytest

from hfcode_amd import fock, molecule as M
from tests.golden_io import CASES, Golden, molecule


@pytest.mark.gpu
@pytest.mark.parametrize("name,mol,basis", [("water_def2svp", M.water, "def2-sv(p)"),
                                            ("benzene_631gd", M.benzene, "6-31g(d)")])
def test_builds_from_reference_texts(name, mol, basis):
    g = Golden(name)
    m = fock.InputMolecule(mol().to_inp(), M.basis_to_dat(M.load_basis(basis)))
    fb = fock.FockBuilder(m)
    S, T, V = fb.onee(m.charges, m.coords_bohr)
    assert np.abs(S - g["S"]).max() < 1e-12 and np.abs(T + V - g.H).max() < 1e-10
    z = np.zeros((m.nb, m.nb))
    err = {"init": np.abs(fb.initial_diag(z, g.P, 5e-7) - g["F_init_5e-7"]).max(),
           "diag": np.abs(fb.diag(z, g.P, 1e-10) - g["F_diag_1e-10"]).max(),
           "nd": np.abs(fb.non_diag(z, g.P, 1e-10) - g["F_nd_1e-10"]).max()}
    print(name, err, flush=True)
    assert max(err.values()) < 1e-10
    fb.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_device_one_electron_matches_reference(name):
    """hfx_ctx_onee's kernels against the reference's S, T, V and the host code."""
    g = Golden(name)
    mol = molecule(name)
    fb = fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    q = mol.znuc.astype(np.float64)
    S, T, V = fb.onee(q, mol.coords_bohr)
    err = [np.abs(S - g["S"]).max(), np.abs(T - g["T"]).max(), np.abs(V - g["V"]).max()]
    print(name, err, flush=True)
    assert err[0] < 1e-13 and err[1] < 1e-13 and err[2] < 1e-12
    for dev, host in zip((S, T, V), fb.onee(q, mol.coords_bohr, device=False)):
        assert np.abs(dev - host).max() < 1e-11
        assert np.array_equal(dev, dev.T)
    fb.close()
