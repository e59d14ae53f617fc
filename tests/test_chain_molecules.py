"""Geometry of the batched / serving workload generator (host only)."""
import numpy as np

from hfcode_amd import molecule as M


def _dmin(x):
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    np.fill_diagonal(d, 9.0)
    return d.min()


def test_chain_formulas_and_bonds():
    for n in (1, 2, 3, 7, 20):
        m = M.chain(n)
        assert m.labels.count("c") == n and m.labels.count("h") == 2 * n + 2
        assert abs(_dmin(m.coords_ang) - 1.09) < 1e-9           # C-H is the shortest contact
    m = M.chain(5, "alkanol")
    assert m.labels.count("o") == 1 and m.labels.count("h") == 2 * 5 and m.nocc * 2 == m.znuc.sum()
    m = M.chain(5, "amine")
    assert m.labels.count("n") == 1 and m.labels.count("h") == 2 * 5 + 1


def test_tetrahedral_angles():
    x = M.chain(4).coords_ang
    c1 = x[1]
    nb = [i for i in range(len(x)) if i != 1 and np.linalg.norm(x[i] - c1) < 1.6]
    assert len(nb) == 4
    for i in nb:
        for j in nb:
            if i < j:
                u, v = x[i] - c1, x[j] - c1
                ang = np.degrees(np.arccos(u @ v / np.linalg.norm(u) / np.linalg.norm(v)))
                assert abs(ang - 109.4712) < 1e-3


def test_batch_and_serving_sets():
    mols = M.batch_molecules()
    b = M.load_basis("def2-sv(p)")
    nbs = [M.ShellTable(m, b).nb for m in mols]
    assert len(mols) == 512 and min(nbs) >= 40 and max(nbs) <= 364
    reqs = M.serving_requests(duration=10.0)
    t = [r[0] for r in reqs]
    assert t == sorted(t) and 100 < len(reqs) < 320
    assert all(5 <= r[1].natom <= 125 for r in reqs)
