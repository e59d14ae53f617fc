"""Batch contexts on the CPU emulation build of libhfx: several molecules of
different sizes and bases in one context, every class launched once for all
of them.

* Each molecule's builds in the batch equal its builds alone (the stand-in
  runs serially, so bit for bit), Schwarz tables included.
* Builds captured in a HIP graph (hfx_batch_fock_device with
  HFX_BUILD_GRAPH; the stand-in records and replays the launches) equal the
  uncaptured builds on every replay, with new densities and thresholds in the
  same device arrays between replays; a molecule with an infinite threshold
  gets nothing.
* The batched SCF (hfx_batch_scf_*) reproduces every molecule's standalone
  SCF (hfx_scf_run): same iteration count, energy within 1e-12 Eh, with and
  without graphs; in mixed precision within 1 uEh of FP64.
"""
import ctypes

import numpy as np
import pytest

from hfcode_amd import molecule as M
from hfcode_amd import scf


def _mols():
    return [(M.water(), "sto-3g"),
            (M.chain(3, "alkanol", jitter=0.05, rng=np.random.default_rng(1)), "def2-sv(p)"),
            (M.chain(2, "amine", jitter=0.05, rng=np.random.default_rng(2)), "6-31g(d)"),
            (M.water(), "def2-sv(p)")]


@pytest.fixture(scope="module")
def emu():
    from tests.conftest import EMU_LIB
    import os
    from hfcode_amd import fock
    if not os.path.exists(EMU_LIB):
        from tests.cpuemu import build_cpuemu
        build_cpuemu.build()
    saved = fock._lib
    fock._lib = fock.configure(ctypes.CDLL(EMU_LIB))
    yield fock
    fock._lib = saved


def make_setup(fock):
    """The batch context of _mols() and each molecule's own context, built
    with whichever library `fock` is bound to."""
    mols = _mols()
    sts = [M.ShellTable(m, M.load_basis(b)) for m, b in mols]
    bb = fock.BatchBuilder(sts)
    S, T, V = bb.onee([m for m, _ in mols])
    alone = []
    for (m, b), st in zip(mols, sts):
        fb = fock.FockBuilder(st)
        s1, t1, v1 = fb.onee(m.znuc.astype(np.float64), m.coords_bohr)
        alone.append(dict(fb=fb, S=s1, H=t1 + v1, P=scf.core_guess(s1, t1 + v1, m.nocc), mol=m))
    return dict(mols=mols, sts=sts, bb=bb, S=S, H=T + V, alone=alone)


@pytest.fixture(scope="module")
def setup(emu):
    st = make_setup(emu)
    yield st
    st["bb"].close()


def test_layout_and_onee(setup):
    bb, alone = setup["bb"], setup["alone"]
    assert bb.nmol == 4 and bb.nb == max(bb.nbs)
    for k, a in enumerate(alone):
        assert np.abs(bb.unstack(setup["S"], k) - a["S"]).max() < 1e-14
        assert np.abs(bb.unstack(setup["H"], k) - a["H"]).max() < 1e-12
        n = bb.nbs[k]
        assert not setup["S"][k, n:, :].any() and not setup["S"][k, :, n:].any()   # padding stays zero


def test_batch_builds_equal_alone(setup, tol=0.0):
    """tol 0: bit for bit, as the one-lane stand-in adds in one order; with
    64-lane waves the order of the atomic additions differs between the
    batch and the single-molecule launch."""
    bb, alone = setup["bb"], setup["alone"]
    P = bb.stack([a["P"] for a in alone])
    Z = np.zeros_like(P)
    Fi = bb.initial_diag(Z, P, 5e-7)
    Fn = bb.non_diag(Z, P, 1e-10)
    Fd = bb.diag(Z, P, 1e-10)
    Fj = bb.non_diag(Z, P, 1e-10, mode="j")
    for k, a in enumerate(alone):
        z = np.zeros_like(a["P"])
        fb = a["fb"]
        assert np.abs(bb.unstack(Fi, k) - fb.initial_diag(z, a["P"], 5e-7)).max() <= tol
        assert np.abs(bb.unstack(Fn, k) - fb.non_diag(z, a["P"], 1e-10)).max() <= tol
        assert np.abs(bb.unstack(Fd, k) - fb.diag(z, a["P"], 1e-10)).max() <= tol
        assert np.abs(bb.unstack(Fj, k) - fb.non_diag(z, a["P"], 1e-10, mode="j")).max() <= tol
        for x, y in zip(bb.schwarz(k), fb.schwarz()):
            assert np.array_equal(x[1], y[1]) and np.array_equal(x[2], y[2]) and x[3] == y[3]
        n = bb.nbs[k]
        assert not Fn[k, n:, :].any() and not Fn[k, :, n:].any()


def test_graph_replays_equal_uncaptured_builds(setup):
    """Device arrays of the stand-in are host memory: the build reads P and
    the thresholds through their addresses, so new values written there
    between replays are what a replay uses."""
    bb, alone = setup["bb"], setup["alone"]
    n = bb.nmol
    P = bb.stack([a["P"] for a in alone])
    bb.initial_diag(np.zeros_like(P), P, 5e-7)
    for a in alone:   # the same Schwarz tables in each molecule's own context (whatever ran before)
        a["fb"].initial_diag(np.zeros_like(a["P"]), a["P"], 5e-7)
    Pd = np.zeros_like(P)
    Fg = np.zeros_like(P)
    Fe = np.zeros_like(P)
    thr = np.zeros(3 * n)
    rng = np.random.default_rng(5)
    for rep in range(3):
        Pd[...] = P * (1.0 + 0.1 * rep)
        thr[:n] = [1e-10, 5e-7, 1e-10, 1e-11][:n]           # diag
        thr[n:2 * n] = [1e-10, 5e-8, np.inf, 1e-11][:n]     # non-diag; molecule 2: none
        thr[2 * n:] = 0.0
        if rep == 2:
            thr[:2 * n] = rng.choice([1e-10, 1e-9, 5e-7], 2 * n)
        for graph, F in ((True, Fg), (False, Fe)):
            F[...] = 7.0      # cleared by the build (HFX_BUILD_ZERO_F)
            bb.build_device(Pd.ctypes.data, F.ctypes.data, thr.ctypes.data, thr[n:].ctypes.data, None,
                            zero_f=True, graph=graph)
        assert np.array_equal(Fg, Fe), rep
        # per molecule: diag at thr_diag + non-diag at thr_nd, as the single-molecule builds
        for k, a in enumerate(alone):
            z = np.zeros_like(a["P"])
            pk = bb.unstack(Pd, k)
            ref = a["fb"].diag(z, pk, thr[k])
            if np.isfinite(thr[n + k]):
                ref = a["fb"].non_diag(ref, pk, thr[n + k])
            assert np.abs(bb.unstack(Fe, k) - ref).max() < 1e-13, (rep, k)


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_scf_equals_alone(setup, graphs):
    bb, alone, mols = setup["bb"], setup["alone"], setup["mols"]
    res = bb.scf(setup["S"], setup["H"], [m.nocc for m, _ in mols], [m.nuclear_repulsion() for m, _ in mols],
                 maxiter=50, diis_start=float("inf"), use_graphs=graphs)
    for k, a in enumerate(alone):
        m = a["mol"]
        dev = a["fb"].scf(a["S"], a["H"], m.nocc, m.nuclear_repulsion(), maxiter=50, diis_start=float("inf"))
        assert res[k].converged and dev.converged
        assert res[k].iterations == dev.iterations
        assert abs(res[k].energy - dev.energy) < 1e-12
        assert np.abs(res[k].P - dev.P).max() < 1e-9
        assert np.abs(res[k].eigs - dev.eigs).max() < 1e-9


def test_batch_scf_mixed_precision(emu, setup):
    mols, sts = setup["mols"], setup["sts"]
    bb = emu.BatchBuilder(sts)
    try:
        bb.set_mixed(1e-5)
        res = bb.scf(setup["S"], setup["H"], [m.nocc for m, _ in mols], [m.nuclear_repulsion() for m, _ in mols],
                     maxiter=50, diis_start=float("inf"))
    finally:
        bb.close()
    ref = setup["bb"].scf(setup["S"], setup["H"], [m.nocc for m, _ in mols],
                          [m.nuclear_repulsion() for m, _ in mols], maxiter=50, diis_start=float("inf"))
    for r, f in zip(res, ref):
        assert r.converged and abs(r.energy - f.energy) < 1e-6


def test_batch_scf_arguments(emu, setup):
    bb = setup["bb"]
    with pytest.raises(emu.HfxError, match="-1"):
        bb.scf(setup["S"], setup["H"], [0, 1, 1, 1], [0.0] * 4)        # no occupied orbitals
    with pytest.raises(emu.HfxError, match="-1"):
        bb.scf(setup["S"], setup["H"], [5, 5, 5, 5], [0.0] * 4, ndiis=17)
