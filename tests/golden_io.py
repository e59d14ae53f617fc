"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["water_sto3g", "benzene_631gd", "water_def2svp", "formaldehyde_631gmix"]
NCOMP = {"s": 1, "p": 3, "d": 5, "l": 4}


def molecule(name):
    """The molecule of a golden case (geometries as in make_golden.py)."""
    from hfcode_amd import molecule as M
    if name.startswith("water"):
        return M.water()
    if name.startswith("benzene"):
        return M.benzene()
    if name.startswith("formaldehyde"):
        return M.Molecule(["c", "o", "h", "h"], [[0, 0, 0], [0, 0, 1.205], [0.9428, 0, -0.5862],
                                                 [-0.9428, 0, -0.5862]], "formaldehyde")
    raise KeyError(name)


class Golden:
    def __init__(self, name):
        self.name = name
        g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.g = g
        self.nb = int(g["nb"])
        self.shells = {k[3:]: g[k] for k in g.files if k.startswith("sh_")}
        self.pairs = [(g["prs_%d" % t], g["schw_%d" % t], g["order_%d" % t], int(g["last_%d" % t]))
                      for t in range(10)]
        self.P = g["P"]
        self.H = g["T"] + g["V"]

    def __getitem__(self, k):
        return self.g[k]

    def ref_block(self, k, cls):
        """Sample k's reference block as an array (n1, n2, n3, n4) in shell order."""
        dims = [NCOMP[c] for c in cls]
        blk = self.g["eri_blocks"][k]
        return blk[:int(np.prod(dims))].reshape(dims[::-1]).transpose(3, 2, 1, 0)

    def desc(self):
        from hfcode_amd import fock
        d, keep, nb, _ = fock.build_desc(self.shells, self.pairs, nb=self.nb)
        self._keep = keep
        return d

    def ntot_pairs(self):
        return sum(len(p[0]) for p in self.pairs)
