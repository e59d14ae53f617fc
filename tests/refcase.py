"""Helpers that set a molecule up in the reference (oracle/_ref/libhfref.so)
and in hfx with the same inputs."""
import numpy as np

import hfref
from hfcode_amd import molecule as M
from tests.conftest import core_guess


class RefCase:
    """Reference set-up of one molecule, a core-guess density, the reference
    Schwarz tables (from fock_build_initial_diag) and the reference builds."""

    def __init__(self, mol, basis_name, thr0=5e-7):
        self.mol = mol
        self.basis = M.load_basis(basis_name)
        self.ref = hfref.RefMolecule.from_molecule(mol, self.basis)
        s, t, v = self.ref.onee()
        self.h = t + v
        self.s = s
        self.nb = self.ref.nb
        self.P = core_guess(s, self.h, self.ref.nocc)
        self.F_init = self.ref.initial_diag(np.zeros((self.nb, self.nb)), self.P, thr0)
        self.shells = self.ref.shells()
        self.pairs = self.ref.pairs()

    def free(self):
        self.ref.free()
