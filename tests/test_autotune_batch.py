"""The batched workload's launch-shape autotuner (hfcode_amd/gen/autotune_batch.py):
its selection from measurements, and (CPU emulation build) that a launch
shape changes no result."""
import numpy as np

from hfcode_amd import molecule as M
from hfcode_amd import scf
from hfcode_amd.gen import autotune_batch as A


def test_select_picks_fastest_per_class():
    data = {"per_cu": {"dlll": {"2": 5.0, "4": 3.0, "8": 4.0, "16": 6.0},
                       "ssss": {"2": 1.0, "4": 1.0, "8": 0.5, "16": 0.9}},
            "chunk": {"dlll": {"16384": 9.0, "65536": 3.1, "262144": 3.5, "0": 3.3},
                      "llll": {"16384": 2.0, "65536": 2.01, "262144": 2.02, "0": 2.0},
                      "ssss": {"16384": 9.0, "65536": 1.0, "262144": 1.0, "0": 1.0}},
            "group": {"16": 5.0, "32": 7.5, "64": 7.4, "128": 6.0}}
    t = A.select(data)
    assert t["classes"]["dlll"] == {"per_cu": 4, "chunk": 65536}
    assert t["classes"]["ssss"] == {"per_cu": 8, "chunk": 0}      # fused class: no chunks
    assert t["classes"]["llll"] == {"per_cu": 0, "chunk": 0}      # chunk times within 2 %: no limit
    assert t["group"] == 32


def test_select_ties_prefer_narrower_launch_and_smaller_group():
    data = {"per_cu": {"sddd": {"2": 1.0, "4": 1.0, "8": 1.0, "16": 1.0}}, "chunk": {},
            "group": {"32": 7.0, "64": 7.0}}
    t = A.select(data)
    assert t["classes"]["sddd"]["per_cu"] == 2 and t["group"] == 32


def test_launch_shape_changes_no_result(emu_fock):
    mols = [(M.water(), "sto-3g"), (M.chain(3, "alkanol", jitter=0.05, rng=np.random.default_rng(1)), "def2-sv(p)"),
            (M.chain(2, "amine", jitter=0.05, rng=np.random.default_rng(2)), "6-31g(d)")]
    sts = [M.ShellTable(m, M.load_basis(b)) for m, b in mols]
    out = []
    for shaped in (False, True):
        bb = emu_fock.BatchBuilder(sts, tuned=False)
        if shaped:
            for ci in range(len(emu_fock.CLASSES)):
                bb.set_class_launch(ci, 2, 256)         # narrowest launch, smallest chunk
        S, T, V = bb.onee([m for m, _ in mols])
        P = bb.stack([scf.core_guess(bb.unstack(S, k), bb.unstack(T + V, k), m.nocc) for k, (m, _) in enumerate(mols)])
        Z = np.zeros_like(P)
        bb.initial_diag(Z, P, 5e-7)
        out.append((bb.non_diag(Z, P, 1e-10), bb.diag(Z, P, 1e-10)))
        bb.close()
    # chunk boundaries move the accumulation order of X: equal up to rounding
    assert np.abs(out[0][0] - out[1][0]).max() < 1e-12 and np.abs(out[0][1] - out[1][1]).max() < 1e-12
