"""Python mirror of the reference's Fock-build interface over libhfx.so.

The reference routines (all FP64, F accumulated in place):

* ``fock_build_non_diag(f, p, thr)``       fock_build_non_diag.f90:1-10
* ``fock_build_non_diag_j_only(f, p, thr)`` fock_build_non_diag_j.f90:1
* ``fock_build_diag(f, p, thr)``           fock_build_diag.f90:1
* ``fock_build_initial_diag(f, p, thr)``   fock_build_initial_diag.f90:1

:class:`FockBuilder` offers the same four operations with the same argument
meaning (``thr`` is the density-weighted Schwarz threshold, ``f`` is added
to, never overwritten).  Every call runs on the GPU through the C ABI of
``include/hfx.h``; there is no CPU fallback: a missing or unloadable
libhfx.so raises.
"""
import ctypes
import os
from types import SimpleNamespace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhfx.so")

PAIR_TYPES = ["ss", "sp", "sd", "sl", "pp", "pd", "pl", "dd", "dl", "ll"]
CLASSES = ("ssss sssp sssd sssl sspp sspd sspl ssdd ssdl ssll spsp spsd spsl sppp sppd "
           "sppl spdd spll sdsd sdsl sdpp sdpd sddd sddl sdll slsl slpp slpl sldd sldl "
           "slll pppp pppd pppl ppdd ppll pdpd pddd plpl plll dddd dddl ddll dldl dlll llll").split()
NTYPES = 10
NCLASSES = 46
HFX_JK, HFX_J_ONLY = 0, 1
OPT_MIXED_SPLIT, OPT_DEBUG_BLOCKS, OPT_PROFILE, OPT_CHUNK, OPT_DIGEST, OPT_OVERLAP = 1, 2, 3, 4, 5, 6
ERRORS = {-1: "bad argument", -2: "unsupported basis", -3: "HIP runtime error",
          -4: "device out of memory", -5: "Schwarz tables missing (run initial_diag first)"}

_LETTER = {1: "s", 2: "p", 3: "d", 12: "l"}


class HfxError(RuntimeError):
    pass


class MolDesc(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int), ("nsh", ctypes.c_int), ("maxprim", ctypes.c_int),
        ("shtyp", ctypes.POINTER(ctypes.c_int)), ("nzet", ctypes.POINTER(ctypes.c_int)),
        ("aosh", ctypes.POINTER(ctypes.c_int)), ("xsh", ctypes.POINTER(ctypes.c_double)),
        ("zeta", ctypes.POINTER(ctypes.c_double)), ("djk", ctypes.POINTER(ctypes.c_double)),
        ("djkl", ctypes.POINTER(ctypes.c_double)),
        ("npair", ctypes.c_int * NTYPES),
        ("prs", ctypes.POINTER(ctypes.c_int) * NTYPES),
        ("schw", ctypes.POINTER(ctypes.c_double) * NTYPES),
        ("order", ctypes.POINTER(ctypes.c_int) * NTYPES),
        ("last", ctypes.c_int * NTYPES),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("quartets", (ctypes.c_longlong * 2) * NCLASSES),
        ("prim_quartets", (ctypes.c_longlong * 2) * NCLASSES),
        ("bra_prims", (ctypes.c_longlong * 2) * NCLASSES),
        ("flops_fp64", ctypes.c_double), ("flops_fp32", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("class_flops", (ctypes.c_double * 2) * NCLASSES),
        ("class_ms", (ctypes.c_float * 2) * NCLASSES),
        ("bra_runs", (ctypes.c_longlong * 2) * NCLASSES),
    ]


class ScfOptions(ctypes.Structure):
    _fields_ = [("maxiter", ctypes.c_int), ("ndiis", ctypes.c_int), ("conv_e", ctypes.c_double),
                ("conv_p", ctypes.c_double), ("lowthresh", ctypes.c_double), ("intthresh", ctypes.c_double),
                ("fdifffac", ctypes.c_double), ("diis_start", ctypes.c_double), ("levelshift", ctypes.c_double),
                ("mode", ctypes.c_int), ("use_graphs", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("fock_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


_lib = None


def release_cached_memory():
    """Hand the device memory that destroyed contexts left cached back to HIP."""
    lib().hfx_release_cached_memory()


def lib():
    """Load libhfx.so; raises when it is missing (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HfxError("libhfx.so not built: run `python -m hfcode_amd.build`")
        _init_torch_first()
        _lib = configure(ctypes.CDLL(LIB_PATH))
    return _lib


def _init_torch_first():
    """PyTorch-ROCm carries its own copy of the HIP runtime, libhfx uses the
    system one; both share the HSA runtime, and torch only finds the GPU when
    its runtime initialises first.  Device tensors handed to the *_device
    entry points come from torch, so torch is initialised before libhfx."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.device_count() > 0:
            torch.cuda.init()
    except Exception:
        pass


def configure(L):
    """Declare the C-ABI signatures of include/hfx.h on a loaded library."""
    dp, ip, vp = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
    L.hfx_ctx_create.argtypes = [ctypes.POINTER(MolDesc), ctypes.POINTER(vp)]
    L.hfx_ctx_destroy.argtypes = [vp]
    L.hfx_ctx_destroy.restype = None
    L.hfx_release_cached_memory.argtypes = []
    L.hfx_release_cached_memory.restype = None
    L.hfx_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.hfx_set_class_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_longlong]
    for n in ("hfx_fock_nd",):
        getattr(L, n).argtypes = [vp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    for n in ("hfx_fock_diag", "hfx_fock_initial_diag"):
        getattr(L, n).argtypes = [vp, dp, dp, ctypes.c_int, ctypes.c_double]
    L.hfx_fock_nd_device.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int, vp]
    L.hfx_fock_diag_device.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    L.hfx_fock_initial_diag_device.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    L.hfx_ctx_stream.argtypes = [vp]
    L.hfx_ctx_stream.restype = vp
    L.hfx_ctx_nb.argtypes = [vp]
    L.hfx_ctx_nmol.argtypes = [vp]
    L.hfx_batch_create.argtypes = [ctypes.POINTER(MolDesc), ctypes.c_int, ctypes.POINTER(vp)]
    L.hfx_batch_get_schwarz.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ip, ip]
    L.hfx_batch_fock_device.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    L.hfx_batch_onee.argtypes = [vp, ip, dp, dp, dp, dp, dp]
    L.hfx_batch_scf_create.argtypes = [vp, dp, dp, ip, dp, dp, ctypes.POINTER(ScfOptions), ctypes.c_int,
                                       ctypes.POINTER(vp)]
    L.hfx_batch_scf_step.argtypes = [vp, ip, dp, ip]
    L.hfx_batch_scf_result.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.hfx_batch_scf_destroy.argtypes = [vp]
    L.hfx_batch_scf_destroy.restype = None
    L.hfx_scf_defaults.argtypes = [ctypes.POINTER(ScfOptions)]
    L.hfx_scf_defaults.restype = None
    L.hfx_scf_run.argtypes = [vp, dp, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.POINTER(ScfOptions),
                              ctypes.POINTER(ScfResult), dp, dp, dp, dp]
    L.hfx_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.hfx_get_schwarz.argtypes = [vp, ctypes.c_int, dp, ip, ip]
    L.hfx_debug_blocks.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip, dp]
    L.hfx_debug_blocks.restype = ctypes.c_longlong
    L.hfx_pair_info.argtypes = [vp, ctypes.c_int, ip, ip]
    L.hfx_onee.argtypes = [ctypes.POINTER(MolDesc), ctypes.c_int, dp, dp, dp, dp, dp]
    L.hfx_ctx_onee.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp, dp]
    L.hfx_ctx_onee_noalloc.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp, dp]
    L.hfx_molecule_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
    L.hfx_molecule_dims.argtypes = [vp, ip, ip, ip, ip, dp]
    L.hfx_molecule_options.argtypes = [vp, ip, ip, dp, ip]
    L.hfx_molecule_desc.argtypes = [vp, ctypes.POINTER(MolDesc)]
    L.hfx_molecule_shells.argtypes = [vp, ip]
    L.hfx_molecule_atoms.argtypes = [vp, dp, dp]
    L.hfx_molecule_density_guess.argtypes = [vp, dp]
    L.hfx_molecule_free.argtypes = [vp]
    L.hfx_molecule_free.restype = None
    L.hfx_version.restype = ctypes.c_char_p
    return L


def _check(rc):
    if rc != 0:
        raise HfxError("hfx error %d: %s" % (rc, ERRORS.get(rc, "unknown")))


def pair_lists(shtyp):
    """Shell-pair lists per type, built as precalc_kabs.f90:38-90 does:
    sh1 = 1..nsh, sh2 = 1..sh1, each pair stored with the shell of the
    earlier letter (s, p, d, L) first.  1-based shells, arrays (n, 2)."""
    order = {1: 0, 2: 1, 3: 2, 12: 3}
    lists = [[] for _ in range(NTYPES)]
    for s1 in range(1, len(shtyp) + 1):
        for s2 in range(1, s1 + 1):
            t1, t2 = int(shtyp[s1 - 1]), int(shtyp[s2 - 1])
            a, b = (s1, s2) if order[t1] <= order[t2] else (s2, s1)
            ta, tb = (t1, t2) if order[t1] <= order[t2] else (t2, t1)
            lists[PAIR_TYPES.index(_LETTER[ta] + _LETTER[tb])].append((a, b))
    return [np.array(l, dtype=np.int32).reshape(-1, 2) for l in lists]


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def build_desc(shells, pairs=None, nb=None):
    """hfx_mol_desc for `shells` (molprops-style arrays) and optional pair
    lists / Schwarz tables; returns (desc, keep-alive list, nb, pair lists)."""
    g = (lambda k: shells[k]) if isinstance(shells, dict) else (lambda k: getattr(shells, k))
    keep = []
    shtyp = np.ascontiguousarray(g("shtyp"), dtype=np.int32)
    nsh = len(shtyp)
    zeta = np.ascontiguousarray(g("zeta"), dtype=np.float64)
    maxprim = zeta.shape[1]
    arrs = dict(
        shtyp=shtyp,
        nzet=np.ascontiguousarray(g("nzet"), dtype=np.int32),
        aosh=np.ascontiguousarray(g("aosh"), dtype=np.int32),
        xsh=np.ascontiguousarray(g("xsh"), dtype=np.float64).reshape(nsh, 3),
        zeta=zeta,
        djk=np.ascontiguousarray(g("djk"), dtype=np.float64),
        djkl=np.ascontiguousarray(g("djkl"), dtype=np.float64),
    )
    keep.append(arrs)
    ncomp = {1: 1, 2: 3, 3: 5, 12: 4}
    nb = int(nb) if nb is not None else int(max(a - 1 + ncomp[int(t)] for a, t in zip(arrs["aosh"], shtyp)))
    d = MolDesc()
    d.nb, d.nsh, d.maxprim = nb, nsh, maxprim
    d.shtyp, d.nzet, d.aosh = _ip(arrs["shtyp"]), _ip(arrs["nzet"]), _ip(arrs["aosh"])
    d.xsh, d.zeta, d.djk, d.djkl = _dp(arrs["xsh"]), _dp(arrs["zeta"]), _dp(arrs["djk"]), _dp(arrs["djkl"])
    if pairs is None:
        pairs = [(p, None, None, 0) for p in pair_lists(shtyp)]
    lists = []
    for t, (prs, schw, order, last) in enumerate(pairs):
        prs = np.ascontiguousarray(prs, dtype=np.int32).reshape(-1, 2)
        lists.append(prs)
        d.npair[t] = len(prs)
        if len(prs) == 0:
            continue
        keep.append(prs)
        d.prs[t] = _ip(prs)
        if schw is not None:
            s = np.ascontiguousarray(schw, dtype=np.float64)
            o = np.ascontiguousarray(order, dtype=np.int32)
            keep += [s, o]
            d.schw[t], d.order[t], d.last[t] = _dp(s), _ip(o), int(last)
    return d, keep, nb, lists


class InputMolecule:
    """A molecule set up by the library's native readers from the texts of a
    reference input file and basis file (``hfx_molecule_read``: read_infile,
    read_basfile with its normalisation, spread_basis).  Carries the shell
    arrays :class:`FockBuilder` takes, the atoms for :meth:`FockBuilder.onee`,
    E_nuc, nocc, the input's SCF options and the atomic-density guess."""

    def __init__(self, input_text, basis_text):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.hfx_molecule_read(input_text.encode(), basis_text.encode(), ctypes.byref(h)))
        self._h = h
        natom, nb, nsh, nocc, unuc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(L.hfx_molecule_dims(h, ctypes.byref(natom), ctypes.byref(nb), ctypes.byref(nsh), ctypes.byref(nocc),
                                   ctypes.byref(unuc)))
        self.natom, self.nb, self.nsh, self.nocc, self.unuc = natom.value, nb.value, nsh.value, nocc.value, unuc.value
        d = MolDesc()
        _check(L.hfx_molecule_desc(h, ctypes.byref(d)))
        n, m = self.nsh, d.maxprim
        self.shtyp = np.ctypeslib.as_array(d.shtyp, (n,)).copy()
        self.nzet = np.ctypeslib.as_array(d.nzet, (n,)).copy()
        self.aosh = np.ctypeslib.as_array(d.aosh, (n,)).copy()
        self.xsh = np.ctypeslib.as_array(d.xsh, (n, 3)).copy()
        self.zeta = np.ctypeslib.as_array(d.zeta, (n, m)).copy()
        self.djk = np.ctypeslib.as_array(d.djk, (n, m)).copy()
        self.djkl = np.ctypeslib.as_array(d.djkl, (n, m)).copy()
        self.atsh = np.zeros(n, np.int32)
        _check(L.hfx_molecule_shells(h, _ip(self.atsh)))
        self.charges = np.zeros(self.natom)
        self.coords_bohr = np.zeros((self.natom, 3))
        _check(L.hfx_molecule_atoms(h, _dp(self.charges), _dp(self.coords_bohr)))
        densup, lev, prt, shift = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(L.hfx_molecule_options(h, ctypes.byref(densup), ctypes.byref(lev), ctypes.byref(shift),
                                      ctypes.byref(prt)))
        self.densup, self.prtlevl = bool(densup.value), prt.value
        self.levelshift = shift.value if lev.value else None

    def density_guess(self):
        """Superposition of the basis file's atomic densities (nb x nb)."""
        P = np.zeros((self.nb, self.nb))
        _check(lib().hfx_molecule_density_guess(self._h, _dp(P)))
        return P

    def close(self):
        if getattr(self, "_h", None):
            lib().hfx_molecule_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FockBuilder:
    """One molecule on the GPU.

    ``shells``: mapping with ``shtyp, nzet, aosh`` (1-based), ``xsh`` (nsh, 3)
    bohr, ``zeta, djk, djkl`` (nsh, maxprim) -- the molprops arrays
    (molprops.f90:93-116), e.g. ``molecule.ShellTable`` or the reference's own.
    ``pairs``: optional list of 10 tuples ``(prs, schw, order, last)`` per pair
    type (the Schwarz tables of fock_build_initial_diag.f90:31-811); when
    omitted the pair lists are built here and the Schwarz tables by the first
    :meth:`initial_diag`.
    """

    def __init__(self, shells, pairs=None, nb=None):
        L = lib()
        self._desc, self._keep, self.nb, self.pair_lists = build_desc(shells, pairs, nb)
        self.nsh = self._desc.nsh
        h = ctypes.c_void_p()
        _check(L.hfx_ctx_create(ctypes.byref(self._desc), ctypes.byref(h)))
        self._h = h
        self.split = 0.0

    def close(self):
        if getattr(self, "_h", None):
            lib().hfx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def onee(self, charges, coords_bohr, device=True):
        """(S, T, V) for point charges at coords (bohr): on the context's GPU
        (hfx_ctx_onee), or with device=False by the host code (hfx_onee)."""
        nb = self.nb
        ch = np.ascontiguousarray(charges, dtype=np.float64)
        xy = np.ascontiguousarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
        S, T, V = (np.zeros((nb, nb)) for _ in range(3))
        if device:
            _check(lib().hfx_ctx_onee(self._h, len(ch), _dp(ch), _dp(xy), _dp(S), _dp(T), _dp(V)))
        else:
            _check(lib().hfx_onee(ctypes.byref(self._desc), len(ch), _dp(ch), _dp(xy), _dp(S), _dp(T), _dp(V)))
        return S, T, V

    def set_mixed(self, split):
        """Quartets whose Schwarz x density bound is below `split` run in FP32."""
        _check(lib().hfx_set_option(self._h, OPT_MIXED_SPLIT, float(split)))
        self.split = float(split)

    def set_profile(self, on=True):
        """Time every class kernel with HIP events (stats()['class_ms'])."""
        _check(lib().hfx_set_option(self._h, OPT_PROFILE, 1.0 if on else 0.0))

    def set_chunk(self, quartets):
        """Quartets per chunk of the block classes (0: as large as possible)."""
        _check(lib().hfx_set_option(self._h, OPT_CHUNK, float(quartets)))

    def set_debug_blocks(self, on=True):
        _check(lib().hfx_set_option(self._h, OPT_DEBUG_BLOCKS, 1.0 if on else 0.0))

    def set_one_pass_digest(self, mode=1):
        """Block classes' digestion: 1 (or True) one pass over each integral
        block, rows of X and P in LDS, where it measured faster and bra runs
        are long (default); 2 wherever it fits; 0 (or False) multi-pass."""
        _check(lib().hfx_set_option(self._h, OPT_DIGEST, float(int(mode))))

    def set_overlap(self, on=True):
        """Block-class lists of a build: each list's first integral kernels
        beside the previous list's last digestion (default), or one list
        after another; same results."""
        _check(lib().hfx_set_option(self._h, OPT_OVERLAP, 1.0 if on else 0.0))

    def _prep(self, f, p):
        nb = self.nb
        p = np.ascontiguousarray(p, dtype=np.float64)
        f = np.array(f, dtype=np.float64, copy=True, order="C")
        if p.shape != (nb, nb) or f.shape != (nb, nb):
            raise ValueError("P and F must be (%d, %d)" % (nb, nb))
        return f, p

    def non_diag(self, f, p, thr, mode="jk", fp64=False):
        """F + G_offdiag(P): fock_build_non_diag (mode 'jk') or
        fock_build_non_diag_j_only (mode 'j').  `fp64`: every quartet in FP64
        for this call whatever the mixed split (the SCF's incremental builds)."""
        f, p = self._prep(f, p)
        m = HFX_J_ONLY if mode == "j" else HFX_JK
        L = lib()
        if fp64 and self.split > 0.0:
            _check(L.hfx_set_option(self._h, OPT_MIXED_SPLIT, 0.0))
            try:
                _check(L.hfx_fock_nd(self._h, _dp(p), _dp(f), self.nb, float(thr), m))
            finally:
                _check(L.hfx_set_option(self._h, OPT_MIXED_SPLIT, self.split))
        else:
            _check(L.hfx_fock_nd(self._h, _dp(p), _dp(f), self.nb, float(thr), m))
        return f

    def diag(self, f, p, thr):
        f, p = self._prep(f, p)
        _check(lib().hfx_fock_diag(self._h, _dp(p), _dp(f), self.nb, float(thr)))
        return f

    def initial_diag(self, f, p, thr):
        f, p = self._prep(f, p)
        _check(lib().hfx_fock_initial_diag(self._h, _dp(p), _dp(f), self.nb, float(thr)))
        return f

    def non_diag_device(self, p_ptr, f_ptr, thr, mode="jk", stream=None):
        """F += G on device arrays (raw device pointers, nb*nb doubles)."""
        m = HFX_J_ONLY if mode == "j" else HFX_JK
        _check(lib().hfx_fock_nd_device(self._h, ctypes.c_void_p(p_ptr), ctypes.c_void_p(f_ptr), float(thr), m,
                                        ctypes.c_void_p(stream or 0)))

    def scf(self, S, H, nocc, unuc, P0=None, levelshift=None, maxiter=25, ndiis=8, conv_e=5e-8, conv_p=1e-5,
            lowthresh=5e-7, intthresh=1e-10, fdifffac=0.1, diis_start=0.3, mode="jk", use_graphs=True):
        """Direct RHF with every step on the GPU (hfx_scf_run): the driver of
        direct_hartreefock.f90:98-223 as hfcode_amd.scf.rhf restates it.
        Returns an object with energy, P, F, eigs, iterations, converged,
        history [(it, etot, deltae, deltap)], t_fock (s) and t_total (s)."""
        nb = self.nb
        S = np.ascontiguousarray(S, dtype=np.float64)
        H = np.ascontiguousarray(H, dtype=np.float64)
        if S.shape != (nb, nb) or H.shape != (nb, nb):
            raise ValueError("S and H must be (%d, %d)" % (nb, nb))
        p0 = None if P0 is None else np.ascontiguousarray(P0, dtype=np.float64)
        o = ScfOptions()
        lib().hfx_scf_defaults(ctypes.byref(o))
        o.maxiter, o.ndiis, o.conv_e, o.conv_p = int(maxiter), int(ndiis), float(conv_e), float(conv_p)
        o.lowthresh, o.intthresh, o.fdifffac = float(lowthresh), float(intthresh), float(fdifffac)
        o.diis_start = min(float(diis_start), 1e300)
        o.levelshift = float(levelshift or 0.0)
        o.mode = HFX_J_ONLY if mode == "j" else HFX_JK
        o.use_graphs = int(bool(use_graphs))
        r = ScfResult()
        P, F, eigs = np.zeros((nb, nb)), np.zeros((nb, nb)), np.zeros(nb)
        hist = np.zeros((maxiter, 4))
        _check(lib().hfx_scf_run(self._h, _dp(S), _dp(H), int(nocc), float(unuc),
                                 _dp(p0) if p0 is not None else None, ctypes.byref(o), ctypes.byref(r),
                                 _dp(P), _dp(F), _dp(eigs), _dp(hist)))
        n = r.iterations
        history = [(i + 1, hist[i, 0], hist[i, 1], hist[i, 2]) for i in range(n)]
        return SimpleNamespace(energy=r.energy, P=P, F=F, eigs=eigs, iterations=n, converged=bool(r.converged),
                               history=history, resmax=hist[:n, 3].copy(), t_fock=r.fock_ms * 1e-3,
                               t_total=r.total_ms * 1e-3)

    def diag_device(self, p_ptr, f_ptr, thr, stream=None):
        _check(lib().hfx_fock_diag_device(self._h, ctypes.c_void_p(p_ptr), ctypes.c_void_p(f_ptr), float(thr),
                                          ctypes.c_void_p(stream or 0)))

    def stats(self):
        return _stats(self._h)

    def schwarz(self):
        out = []
        for t, prs in enumerate(self.pair_lists):
            n = len(prs)
            s = np.zeros(max(n, 1)); o = np.zeros(max(n, 1), np.int32); last = ctypes.c_int(0)
            _check(lib().hfx_get_schwarz(self._h, t, _dp(s), _ip(o), ctypes.byref(last)))
            out.append((prs, s[:n], o[:n], last.value))
        return out

    def debug_blocks(self, cls, prec=0):
        ci = CLASSES.index(cls) if isinstance(cls, str) else int(cls)
        n = lib().hfx_debug_blocks(self._h, ci, prec, None, None)
        if n < 0:
            _check(int(n))
        if n == 0:
            return np.zeros((0, 2), np.int32), np.zeros((0, 625))
        q = np.zeros((n, 2), np.int32)
        b = np.zeros((n, 625))
        lib().hfx_debug_blocks(self._h, ci, prec, _ip(q), _dp(b))
        return q, b

    def pair_info(self, g):
        t, i = ctypes.c_int(), ctypes.c_int()
        _check(lib().hfx_pair_info(self._h, int(g), ctypes.byref(t), ctypes.byref(i)))
        return t.value, i.value


BUILD_ZERO_F, BUILD_GRAPH = 1, 2


def _stats(h):
    """hfx_get_stats of a context as a dict of arrays."""
    s = Stats()
    _check(lib().hfx_get_stats(h, ctypes.byref(s)))
    q = np.array([[s.quartets[c][k] for k in range(2)] for c in range(NCLASSES)], dtype=np.int64)
    pq = np.array([[s.prim_quartets[c][k] for k in range(2)] for c in range(NCLASSES)], dtype=np.int64)
    bp = np.array([[s.bra_prims[c][k] for k in range(2)] for c in range(NCLASSES)], dtype=np.int64)
    cf = np.array([[s.class_flops[c][k] for k in range(2)] for c in range(NCLASSES)])
    cm = np.array([[s.class_ms[c][k] for k in range(2)] for c in range(NCLASSES)])
    br = np.array([[s.bra_runs[c][k] for k in range(2)] for c in range(NCLASSES)], dtype=np.int64)
    return dict(quartets=q, prim_quartets=pq, bra_prims=bp, flops_fp64=s.flops_fp64,
                flops_fp32=s.flops_fp32, ms=s.ms_total, class_flops=cf, class_ms=cm, bra_runs=br)


class BatchBuilder:
    """Several molecules in one context (``hfx_batch_create``): each quartet
    class is screened and evaluated for all of them in one launch of its
    kernels.  Matrices are stacked ``(nmol, nb, nb)`` arrays with ``nb`` the
    largest molecule's basis size; molecule m uses ``[m, :nb_m, :nb_m]``
    (:meth:`stack` / :meth:`unstack`).  The Fock-build methods take and
    return such stacks, with the meaning of :class:`FockBuilder`'s."""

    def __init__(self, shell_tables, tuned=True):
        """`tuned`: apply the launch shapes of hfcode_amd/gen/batch_launch.json
        (autotune_batch.py) when that table exists."""
        L = lib()
        n = len(shell_tables)
        if n < 1:
            raise ValueError("a batch needs at least one molecule")
        self._descs = (MolDesc * n)()
        self._keep = []
        self.nbs, self.pair_lists = [], []
        for i, st in enumerate(shell_tables):
            d, keep, nb, lists = build_desc(st)
            self._descs[i] = d
            self._keep.append(keep)
            self.nbs.append(nb)
            self.pair_lists.append(lists)
        h = ctypes.c_void_p()
        _check(L.hfx_batch_create(self._descs, n, ctypes.byref(h)))
        self._h = h
        self.nmol = n
        self.nb = L.hfx_ctx_nb(h)
        self.split = 0.0
        if tuned:
            from .gen import autotune_batch
            tab = autotune_batch.load_table()
            if tab:
                for ci, c in enumerate(CLASSES):
                    e = tab["classes"].get(c)
                    if e:
                        self.set_class_launch(ci, e["per_cu"], e["chunk"])

    def set_class_launch(self, cls, per_cu, chunk):
        """Launch shape of class `cls` (index or name): workgroups per CU and
        quartets per chunk (0: defaults), hfx_set_class_launch."""
        ci = CLASSES.index(cls) if isinstance(cls, str) else int(cls)
        _check(lib().hfx_set_class_launch(self._h, ci, int(per_cu), int(chunk)))

    def set_profile(self, on=True):
        _check(lib().hfx_set_option(self._h, OPT_PROFILE, 1.0 if on else 0.0))

    def stats(self):
        return _stats(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().hfx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stack(self, mats):
        """(nmol, nb, nb) array from per-molecule matrices (zero padding)."""
        out = np.zeros((self.nmol, self.nb, self.nb))
        for m, a in enumerate(mats):
            a = np.asarray(a, dtype=np.float64)
            out[m, :a.shape[0], :a.shape[1]] = a
        return out

    def unstack(self, arr, m):
        n = self.nbs[m]
        return np.array(arr[m, :n, :n])

    def set_mixed(self, split):
        _check(lib().hfx_set_option(self._h, OPT_MIXED_SPLIT, float(split)))
        self.split = float(split)

    def onee(self, mols):
        """Stacked (S, T, V) of the molecules (objects with znuc, coords_bohr)."""
        nat = np.array([len(m.znuc) for m in mols], dtype=np.int32)
        ch = np.ascontiguousarray(np.concatenate([np.asarray(m.znuc, dtype=np.float64) for m in mols]))
        xy = np.ascontiguousarray(np.concatenate([np.asarray(m.coords_bohr, dtype=np.float64).reshape(-1, 3)
                                                  for m in mols]))
        S, T, V = (np.zeros((self.nmol, self.nb, self.nb)) for _ in range(3))
        _check(lib().hfx_batch_onee(self._h, _ip(nat), _dp(ch), _dp(xy), _dp(S), _dp(T), _dp(V)))
        return S, T, V

    def _prep(self, f, p):
        shape = (self.nmol, self.nb, self.nb)
        p = np.ascontiguousarray(p, dtype=np.float64)
        f = np.array(f, dtype=np.float64, copy=True, order="C")
        if p.shape != shape or f.shape != shape:
            raise ValueError("P and F must be %r" % (shape,))
        return f, p

    def non_diag(self, f, p, thr, mode="jk"):
        f, p = self._prep(f, p)
        _check(lib().hfx_fock_nd(self._h, _dp(p), _dp(f), self.nb, float(thr), HFX_J_ONLY if mode == "j" else HFX_JK))
        return f

    def diag(self, f, p, thr):
        f, p = self._prep(f, p)
        _check(lib().hfx_fock_diag(self._h, _dp(p), _dp(f), self.nb, float(thr)))
        return f

    def initial_diag(self, f, p, thr):
        f, p = self._prep(f, p)
        _check(lib().hfx_fock_initial_diag(self._h, _dp(p), _dp(f), self.nb, float(thr)))
        return f

    def build_device(self, p_ptr, f_ptr, thr_diag_ptr, thr_nd_ptr, split_ptr=None, mode="jk", zero_f=True,
                     graph=True, stream=None):
        """hfx_batch_fock_device on device arrays (raw pointers): F (+)= the
        diag + non-diag build of every molecule at its own thresholds."""
        flags = (BUILD_ZERO_F if zero_f else 0) | (BUILD_GRAPH if graph else 0)
        _check(lib().hfx_batch_fock_device(self._h, ctypes.c_void_p(p_ptr), ctypes.c_void_p(f_ptr),
                                           ctypes.c_void_p(thr_diag_ptr), ctypes.c_void_p(thr_nd_ptr),
                                           ctypes.c_void_p(split_ptr or 0), HFX_J_ONLY if mode == "j" else HFX_JK,
                                           flags, ctypes.c_void_p(stream or 0)))

    def schwarz(self, m):
        out = []
        for t, prs in enumerate(self.pair_lists[m]):
            n = len(prs)
            sc = np.zeros(max(n, 1)); o = np.zeros(max(n, 1), np.int32); last = ctypes.c_int(0)
            _check(lib().hfx_batch_get_schwarz(self._h, m, t, _dp(sc), _ip(o), ctypes.byref(last)))
            out.append((prs, sc[:n], o[:n], last.value))
        return out

    def scf(self, S, H, nocc, unuc, P0=None, maxiter=25, ndiis=8, conv_e=5e-8, conv_p=1e-5, lowthresh=5e-7,
            intthresh=1e-10, fdifffac=0.1, diis_start=0.3, mode="jk", use_graphs=True, on_done=None):
        """Direct RHF of every molecule at once on the GPU (hfx_batch_scf_*).
        S, H, P0: stacked arrays; nocc, unuc: one per molecule.  `on_done(m,
        result)` is called as each molecule finishes.  Returns a list of
        SimpleNamespace(energy, P, F, eigs, iterations, converged, status)."""
        S = np.ascontiguousarray(S, dtype=np.float64)
        H = np.ascontiguousarray(H, dtype=np.float64)
        p0 = None if P0 is None else np.ascontiguousarray(P0, dtype=np.float64)
        occ = np.ascontiguousarray(nocc, dtype=np.int32)
        un = np.ascontiguousarray(unuc, dtype=np.float64)
        o = ScfOptions()
        L = lib()
        L.hfx_scf_defaults(ctypes.byref(o))
        o.maxiter, o.ndiis, o.conv_e, o.conv_p = int(maxiter), int(ndiis), float(conv_e), float(conv_p)
        o.lowthresh, o.intthresh, o.fdifffac = float(lowthresh), float(intthresh), float(fdifffac)
        o.diis_start = min(float(diis_start), 1e300)
        o.levelshift = 0.0
        o.mode = HFX_J_ONLY if mode == "j" else HFX_JK
        h = ctypes.c_void_p()
        _check(L.hfx_batch_scf_create(self._h, _dp(S), _dp(H), _ip(occ), _dp(un),
                                      _dp(p0) if p0 is not None else None, ctypes.byref(o), int(bool(use_graphs)),
                                      ctypes.byref(h)))
        n = self.nmol
        status = np.zeros(n, np.int32)
        energy = np.zeros(n)
        iters = np.zeros(n, np.int32)
        results = [None] * n
        try:
            while True:
                live = L.hfx_batch_scf_step(h, _ip(status), _dp(energy), _ip(iters))
                if live < 0:
                    _check(live)
                for m in range(n):
                    if results[m] is None and status[m] != 0:
                        nbm = self.nbs[m]
                        P, F, e = np.zeros((nbm, nbm)), np.zeros((nbm, nbm)), np.zeros(nbm)
                        _check(L.hfx_batch_scf_result(h, m, _dp(P), _dp(F), _dp(e)))
                        results[m] = SimpleNamespace(energy=float(energy[m]), P=P, F=F, eigs=e,
                                                     iterations=int(iters[m]), converged=bool(status[m] == 1),
                                                     status=int(status[m]))
                        if on_done is not None:
                            on_done(m, results[m])
                if live == 0:
                    break
        finally:
            L.hfx_batch_scf_destroy(h)
        return results
