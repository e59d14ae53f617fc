"""libhfx's own sources (host code, support kernels, generated class kernels)
compiled for the CPU against the serial HIP stand-in of tests/cpuemu, checked
against the golden fixtures and the oracle.  This exercises the screening,
work lists, block/fused kernels and digestion logic without a GPU; the GPU
build of the same sources is checked by the -m gpu tests."""
import ctypes

import numpy as np
import pytest

from hfcode_amd import molecule as M
from hfcode_amd import scf
from tests.golden_io import CASES, NCOMP, Golden, molecule

TOL = 1e-10


@pytest.mark.parametrize("name", CASES)
def test_builds_match_golden(emu_fock, name):
    g = Golden(name)
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    z = np.zeros((g.nb, g.nb))
    assert np.abs(fb.initial_diag(z, g.P, 5e-7) - g["F_init_5e-7"]).max() < TOL
    assert np.abs(fb.diag(z, g.P, 1e-10) - g["F_diag_1e-10"]).max() < TOL
    if name == "formaldehyde_631gmix":
        import oracle
        Fo, _ = oracle.fock(g.desc(), g.P, z, 1e-10, "non_diag")
        assert np.abs(fb.non_diag(z, g.P, 1e-10) - Fo).max() < TOL
        return
    assert np.abs(fb.non_diag(z, g.P, 1e-10) - g["F_nd_1e-10"]).max() < TOL
    assert np.abs(fb.non_diag(z, g.P, 5e-7) - g["F_nd_5e-7"]).max() < TOL


@pytest.mark.parametrize("name", ["water_def2svp", "formaldehyde_631gmix"])
def test_class_blocks_match_oracle(emu_fock, name):
    """Every block the generated kernels produced for the build, against the
    oracle's McMurchie--Davidson block on the same four shells."""
    import oracle
    g = Golden(name)
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    fb.set_debug_blocks(True)
    fb.non_diag(np.zeros((g.nb, g.nb)), g.P, 1e-10)
    d = g.desc()
    seen = set()
    for ci, cls in enumerate(emu_fock.CLASSES):
        q, b = fb.debug_blocks(ci)
        for k in range(min(len(q), 12)):
            t1, i1 = fb.pair_info(q[k, 0])
            t2, i2 = fb.pair_info(q[k, 1])
            s1, s2 = g.pairs[t1][0][i1 - 1]
            s3, s4 = g.pairs[t2][0][i2 - 1]
            dims = [NCOMP[c] for c in cls]
            o = oracle.eri_block(d, s1, s2, s3, s4, dims, skip_negab=(cls != "slsl"))
            mine = b[k][:int(np.prod(dims))].reshape(dims[::-1]).transpose(3, 2, 1, 0)
            assert np.abs(mine - o).max() <= 1e-11 * np.abs(o).max() + 1e-15, (cls, s1, s2, s3, s4)
            seen.add(cls)
    assert len(seen) >= 20


def test_fp32_class_blocks_match_oracle(emu_fock):
    """Every quartet routed to FP32 (split above every bound): the FP32
    kernels' blocks (the block classes group several L-part combinations per
    kernel in FP32) against the oracle's FP64 blocks, to single precision."""
    import oracle
    g = Golden("formaldehyde_631gmix")
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    fb.set_mixed(1e30)
    fb.set_debug_blocks(True)
    fb.non_diag(np.zeros((g.nb, g.nb)), g.P, 1e-10)
    d = g.desc()
    seen = set()
    for ci, cls in enumerate(emu_fock.CLASSES):
        q, b = fb.debug_blocks(ci, prec=1)
        for k in range(min(len(q), 6)):
            t1, i1 = fb.pair_info(q[k, 0])
            t2, i2 = fb.pair_info(q[k, 1])
            s1, s2 = g.pairs[t1][0][i1 - 1]
            s3, s4 = g.pairs[t2][0][i2 - 1]
            dims = [NCOMP[c] for c in cls]
            o = oracle.eri_block(d, s1, s2, s3, s4, dims, skip_negab=(cls != "slsl"))
            mine = b[k][:int(np.prod(dims))].reshape(dims[::-1]).transpose(3, 2, 1, 0)
            assert np.abs(mine - o).max() <= 2e-6 * np.abs(o).max() + 1e-12, (cls, s1, s2, s3, s4)
            seen.add(cls)
    assert len(seen) >= 20


def test_j_only_matches_oracle(emu_fock):
    import oracle
    g = Golden("water_def2svp")
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    z = np.zeros((g.nb, g.nb))
    Fo, _ = oracle.fock(g.desc(), g.P, z, 1e-10, "j_only")
    assert np.abs(fb.non_diag(z, g.P, 1e-10, mode="j") - Fo).max() < TOL


def test_mixed_precision_split(emu_fock):
    g = Golden("benzene_631gd")
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    F64 = fb.non_diag(g.H, g.P, 1e-10)
    fb.set_mixed(1e-5)
    F32 = fb.non_diag(g.H, g.P, 1e-10)
    q = fb.stats()["quartets"]
    assert q[:, 1].sum() > 0
    dE = 0.5 * np.sum(g.P * (F32 - F64))
    assert abs(dE) < 1e-8


@pytest.mark.parametrize("name", ["water_def2svp", "benzene_631gd", "formaldehyde_631gmix"])
def test_all_fp32_build_matches_fp64(emu_fock, name):
    """Every quartet through the FP32 kernels and the FP32 digestion (both
    of its K forms occur among the classes): F within FP32 rounding of the
    FP64 build (measured max |dF| 1e-6..3e-6 at max |F| ~ 15; a wrong index
    in a contraction moves F by ~|e| |P| >= 1e-2)."""
    g = Golden(name)
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    z = np.zeros((g.nb, g.nb))
    F64 = fb.non_diag(z, g.P, 1e-10)
    fb.set_mixed(1e30)
    F32 = fb.non_diag(z, g.P, 1e-10)
    q = fb.stats()["quartets"]
    assert q[:, 0].sum() == 0 and q[:, 1].sum() > 0
    assert np.abs(F32 - F64).max() < 2e-5


def test_chunked_builds(emu_fock):
    """The block classes in chunks of 256 quartets.  With one live context the
    chunks alternate between the two halves of the integral buffer (a chunk's
    integrals beside the previous chunk's digestion); with a second context
    live they run one after another.  Both give the golden Fock matrix, and
    the chunked mixed build the unchunked one."""
    import gc
    gc.collect()
    g = Golden("benzene_631gd")
    z = np.zeros((g.nb, g.nb))
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    fb.set_chunk(256)
    one = fb.non_diag(z, g.P, 1e-10)
    assert np.abs(one - g["F_nd_1e-10"]).max() < TOL
    other = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    other.set_chunk(256)
    assert np.abs(other.non_diag(z, g.P, 1e-10) - g["F_nd_1e-10"]).max() < TOL
    other.close()
    fb.set_mixed(1e-5)
    mixed_chunked = fb.non_diag(z, g.P, 1e-10)
    fb.set_chunk(0)
    assert np.abs(mixed_chunked - fb.non_diag(z, g.P, 1e-10)).max() < 1e-12
    fb.close()


@pytest.mark.parametrize("split,chunk", [(0.0, 0), (0.0, 256), (1e30, 0), (1e-5, 512)])
def test_one_pass_digestion_equals_multi_pass(emu_fock, split, chunk):
    """The block classes' one-pass digestion (resident bra's rows of X and P
    in LDS, steps that never cross a bra pair) against the multi-pass one:
    the same sums in another order, so FP64 builds agree to rounding and
    all-FP32 builds to FP32 rounding of the per-quartet sums (a wrong row,
    index or weight moves F by ~|e| |P| >= 1e-2).  Chunks of 256 and 512
    quartets cut bra pairs' runs at chunk and workgroup boundaries."""
    g = Golden("benzene_631gd")
    z = np.zeros((g.nb, g.nb))
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    fb.set_mixed(split)
    fb.set_chunk(chunk)
    fb.set_one_pass_digest(2)   # every block class that fits, whatever the runs
    one = fb.non_diag(z, g.P, 1e-10)
    fb.set_one_pass_digest(False)
    multi = fb.non_diag(z, g.P, 1e-10)
    fb.close()
    tol = 1e-12 if split == 0.0 else 1e-5
    assert np.abs(one - multi).max() < tol
    if split == 0.0:
        assert np.abs(one - g["F_nd_1e-10"]).max() < TOL


@pytest.mark.parametrize("split,chunk", [(0.0, 256), (1e-5, 512)])
def test_overlapping_lists_equal_lists_in_turn(emu_fock, split, chunk):
    """Block-class lists in half-buffer chunks numbered across the build
    (HFX_OPT_OVERLAP 1, the default with one live context) against one list
    after another: the same quartets, added in another order."""
    g = Golden("benzene_631gd")
    z = np.zeros((g.nb, g.nb))
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    fb.set_mixed(split)
    fb.set_chunk(chunk)
    fb.set_overlap(True)
    over = [fb.non_diag(z, g.P, 1e-10) for _ in range(2)]
    fb.set_overlap(False)
    turn = fb.non_diag(z, g.P, 1e-10)
    fb.close()
    for F in over:
        assert np.abs(F - turn).max() < 1e-12
    if split == 0.0:
        assert np.abs(turn - g["F_nd_1e-10"]).max() < TOL


@pytest.mark.parametrize("name", ["water_sto3g", "water_def2svp"])
def test_standalone_path_and_scf(emu_fock, name):
    """Own shells, one-electron matrices, Schwarz tables and SCF driver
    against the reference's S, T, V, Schwarz tables and converged energy."""
    g = Golden(name)
    basis = {"water_sto3g": "sto-3g", "water_def2svp": "def2-sv(p)"}[name]
    mol = M.water()
    fb = emu_fock.FockBuilder(M.ShellTable(mol, M.load_basis(basis)))
    S, T, V = fb.onee(mol.znuc.astype(np.float64), mol.coords_bohr)
    assert np.abs(S - g["S"]).max() < 1e-12
    assert np.abs(T - g["T"]).max() < 1e-11 and np.abs(V - g["V"]).max() < 1e-10
    assert np.allclose(np.diag(S), 1.0, atol=1e-12)          # normalised contractions
    Fi = fb.initial_diag(np.zeros_like(S), g.P, 5e-7)
    assert np.abs(Fi - g["F_init_5e-7"]).max() < TOL
    for (p1, s1, o1, l1), (p2, s2, o2, l2) in zip(g.pairs, fb.schwarz()):
        assert np.array_equal(p1, p2) and l1 == l2
        if len(s1):
            assert np.allclose(s1, s2, rtol=1e-12, atol=1e-15)
    res = scf.rhf(fb, S, T + V, mol.nocc, mol.nuclear_repulsion())
    assert res.converged and abs(res.energy - float(g["E_scf"])) < 1e-9


@pytest.mark.parametrize("name,basis,ls", [("water_sto3g", "sto-3g", None), ("water_def2svp", "def2-sv(p)", None),
                                            ("benzene_631gd", "6-31g(d)", 0.4)])
def test_device_scf_driver(emu_fock, name, basis, ls):
    """hfx_scf_run (the SCF driver's device code, with serial stand-ins for
    dgemm and dsyevd) against the reference's converged energy and the host
    restatement of the driver, iteration by iteration."""
    g = Golden(name)
    mol = molecule(name)
    st = M.ShellTable(mol, M.load_basis(basis))
    fb = emu_fock.FockBuilder(st)
    S, T, V = fb.onee(mol.znuc.astype(np.float64), mol.coords_bohr)
    dev = fb.scf(S, T + V, mol.nocc, mol.nuclear_repulsion(), levelshift=ls)
    host = scf.rhf(emu_fock.FockBuilder(st), S, T + V, mol.nocc, mol.nuclear_repulsion(), levelshift=ls)
    assert dev.converged and abs(dev.energy - float(g["E_scf"])) < 1e-9
    assert dev.iterations == host.iterations
    for a, b in zip(dev.history, host.history):
        assert abs(a[1] - b[1]) < 1e-9 and abs(a[3] - b[3]) < 1e-8, a[0]
    assert np.abs(dev.P - host.P).max() < 1e-9 and np.abs(dev.F - host.F).max() < 1e-9


def test_device_scf_driver_arguments(emu_fock):
    g = Golden("water_sto3g")
    fb = emu_fock.FockBuilder(g.shells, nb=g.nb)
    S = np.eye(g.nb)
    with pytest.raises(emu_fock.HfxError, match="-1"):
        fb.scf(S, S, 0, 0.0)                       # no occupied orbitals
    with pytest.raises(emu_fock.HfxError, match="-1"):
        fb.scf(S, S, 5, 0.0, ndiis=17)             # more DIIS vectors than the driver keeps
    with pytest.raises(ValueError):
        fb.scf(np.eye(3), S, 5, 0.0)


@pytest.mark.parametrize("name", CASES)
def test_device_one_electron_matches_reference(emu_fock, name):
    """Generated S/T/V kernels (hfx_ctx_onee) against the reference's S, T, V
    (int1e / int1e_kin / int1e_nuc matrices of the golden case) and against
    the host code (hfx_onee)."""
    g = Golden(name)
    mol = molecule(name)
    fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
    q = mol.znuc.astype(np.float64)
    S, T, V = fb.onee(q, mol.coords_bohr)
    assert np.abs(S - g["S"]).max() < 1e-13 and np.abs(T - g["T"]).max() < 1e-13
    assert np.abs(V - g["V"]).max() < 1e-12
    for dev, host in zip((S, T, V), fb.onee(q, mol.coords_bohr, device=False)):
        assert np.abs(dev - host).max() < 1e-11
        assert np.array_equal(dev, dev.T)


def test_contexts_reuse_cached_device_memory(emu_fock):
    """Contexts created after others were destroyed (blocks from the cache)
    build the same Fock matrix; releasing the cache leaves later contexts
    working."""
    g = Golden("water_def2svp")
    z = np.zeros((g.nb, g.nb))
    ref = None
    for k in range(3):
        fb = emu_fock.FockBuilder(g.shells, g.pairs, nb=g.nb)
        F = fb.non_diag(z, g.P, 1e-10)
        fb.close()
        if ref is None:
            ref = F
        assert np.array_equal(F, ref)
        if k == 1:
            emu_fock.release_cached_memory()


def test_error_codes(emu_fock):
    f = emu_fock
    g = Golden("water_sto3g")
    d, keep, nb, _ = f.build_desc(g.shells, g.pairs, nb=g.nb)
    h = ctypes.c_void_p()
    L = f.lib()
    assert L.hfx_ctx_create(None, ctypes.byref(h)) == -1
    # P + D + L mix is rejected like spread_basis.f90:174-177
    sh = {k: np.array(v) for k, v in g.shells.items()}
    sh["shtyp"] = sh["shtyp"].copy()
    sh["shtyp"][0] = 2          # O: p, L (kept), H: d  ->  P + D + L
    sh["shtyp"][2] = 3
    with pytest.raises(f.HfxError):
        f.FockBuilder(sh, nb=g.nb + 10)
    fb = f.FockBuilder(g.shells, nb=g.nb)              # no Schwarz tables yet
    with pytest.raises(f.HfxError, match="Schwarz"):
        fb.non_diag(np.zeros((g.nb, g.nb)), g.P, 1e-10)
    with pytest.raises(ValueError):
        fb.diag(np.zeros((3, 3)), g.P, 1e-10)


def test_more_than_ten_primitives_rejected(emu_fock):
    """A shell with more primitives than the reference's maxprim (10,
    molprops.f90:9-18) is refused at context creation: the screening keeps
    the set of survivors' primitive-pair counts + 1 in 128 bits."""
    g = Golden("water_sto3g")
    sh = {k: np.array(v) for k, v in g.shells.items()}
    mp = 16
    for k in ("zeta", "djk", "djkl"):
        a = np.zeros((len(sh["shtyp"]), mp))
        a[:, :sh[k].shape[1]] = sh[k]
        sh[k] = a
    ok = emu_fock.FockBuilder(sh, nb=g.nb)       # a leading dimension of 16 is fine
    ok.close()
    sh["nzet"] = sh["nzet"].copy()
    sh["nzet"][0] = 11
    sh["zeta"][0, :11] = np.geomspace(100.0, 0.1, 11)
    sh["djk"][0, :11] = 0.1
    with pytest.raises(emu_fock.HfxError, match="unsupported"):
        emu_fock.FockBuilder(sh, nb=g.nb)
