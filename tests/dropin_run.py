"""Child process of tests/test_gpu_dropin.py: one molecule through the
reference set-up with fock_build_non_diag replaced by the hfx Fortran shim
(oracle/_ref/libhfref_hip.so).  Prints one JSON line."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hfref  # noqa: E402
from hfcode_amd import molecule as M, scf  # noqa: E402
from tests.golden_io import Golden  # noqa: E402


def main():
    name, molname, basis, ls = sys.argv[1], sys.argv[2], sys.argv[3], sys.argv[4]
    ls = None if ls == "none" else float(ls)
    hfref.use_library(hfref.DROPIN_PATH)
    g = Golden(name)
    r = hfref.RefMolecule.from_molecule(getattr(M, molname)(), M.load_basis(basis))
    s, t, v = r.onee()
    z = np.zeros((r.nb, r.nb))
    Fi = r.initial_diag(z, g.P, 5e-7)                  # reference routine
    out = {"F_init": float(np.abs(Fi - g["F_init_5e-7"]).max())}
    for thr, key in ((5e-7, "F_nd_5e-7"), (1e-10, "F_nd_1e-10")):
        F = r.non_diag(z, g.P, thr)                     # shim -> libhfx on the GPU
        out[key] = float(np.abs(F - g[key]).max())
    res = scf.rhf(r, s, t + v, r.nocc, r.unuc, levelshift=ls)
    out["dE"] = float(res.energy - float(g["E_scf"]))
    out["converged"] = bool(res.converged)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
