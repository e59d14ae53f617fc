"""Molecules, basis sets and the shell table the Fock build consumes.

This is the host-side set-up the drop-in needs when it runs without the
Fortran front end (batched and serving jobs).  It restates, in Python:

* the input conventions of ``read_infile`` (read_infile.f90:12-73): atoms in
  Angstrom, converted with ``bohr = 1.889725989``;
* contraction normalisation of ``normalize`` (read_basfile.f90:100-179);
* the shell / AO ordering of ``spread_basis`` (spread_basis.f90:53-150):
  atoms in input order, shells in basis-file order, AOs numbered shell by
  shell, s = 1 AO, p = 3 (x, y, z), d = 5 spherical, L = 4 (s, x, y, z);
  an L shell keeps its s coefficients in ``djk`` and the p/s ratio in ``djkl``
  (spread_basis.f90:149-151).

It also writes the reference's own ``.inp`` / ``.dat`` text formats, so the
same molecule can be handed to the reference program unchanged.
"""
import json
import math
import os

import numpy as np

BOHR = 1.889725989            # read_infile.f90:7
PI = 3.141592653589793238462643383279502884197
PI32 = math.sqrt(PI ** 3)     # nrtype.f90:11

SHELL_CODE = {"s": 1, "p": 2, "d": 3, "l": 12}   # molprops.f90:93 (shtyp)
SHELL_NFUNC = {1: 1, 2: 3, 3: 5, 12: 4}
ELEMENTS = ["h", "he", "li", "be", "b", "c", "n", "o", "f", "ne",
            "na", "mg", "al", "si", "p", "s", "cl", "ar",
            "k", "ca", "sc", "ti", "v", "cr", "mn", "fe", "co", "ni", "cu", "zn",
            "ga", "ge", "as", "se", "br", "kr"]     # read_infile.f90:126-130

_BASIS_DIR = os.path.join(os.path.dirname(__file__), "basis")
BASIS_FILES = {"sto-3g": "sto-3g.json", "6-31g(d)": "6-31gd.json",
               "def2-sv(p)": "def2-svp_p.json", "6-31g-mix": "6-31g-mix.json"}


def load_basis(name):
    """Basis set by name (``sto-3g``, ``6-31g(d)``, ``def2-sv(p)``, ``6-31g-mix``)."""
    with open(os.path.join(_BASIS_DIR, BASIS_FILES[name.lower()])) as fh:
        return json.load(fh)


class Molecule:
    """Atoms (lower-case element labels) and coordinates in Angstrom."""

    def __init__(self, labels, coords_ang, title="molecule"):
        self.labels = [l.lower() for l in labels]
        self.coords_ang = np.asarray(coords_ang, dtype=np.float64).reshape(-1, 3)
        self.title = title
        for l in self.labels:
            if l not in ELEMENTS:
                raise ValueError("element not recognized: %r" % l)

    @property
    def natom(self):
        return len(self.labels)

    @property
    def znuc(self):
        return np.array([ELEMENTS.index(l) + 1 for l in self.labels], dtype=np.int64)

    @property
    def nocc(self):
        return int(self.znuc.sum()) // 2           # read_infile.f90:78

    @property
    def coords_bohr(self):
        return self.coords_ang * BOHR               # read_infile.f90:70

    def nuclear_repulsion(self):
        """E_nuc, nuclear_repulsion.f90:1-20."""
        x = self.coords_bohr
        z = self.znuc.astype(np.float64)
        e = 0.0
        for i in range(self.natom):
            for j in range(i + 1, self.natom):
                e += z[i] * z[j] / math.sqrt(float(((x[i] - x[j]) ** 2).sum()))
        return e

    def to_inp(self, prtlevl=1, levelshift=None):
        """Text of the reference's current-format input (read_infile.f90:12-61)."""
        ls = "LEVELSHIFT Y %s" % repr(float(levelshift)) if levelshift else "LEVELSHIFT N 0.0"
        lines = ["%d" % self.natom, self.title.replace(" ", "_")[:40],
                 "USEDENSITYSUPGUESS N", "DIRECTSCF Y", ls,
                 "UHF N 0", "MP2 N", "CCSD N", "PRINT %d" % prtlevl, "ncore 0"]
        for l, (x, y, z) in zip(self.labels, self.coords_ang):
            lines.append("%-2s %s %s %s" % (l.capitalize(), repr(float(x)), repr(float(y)), repr(float(z))))
        return "\n".join(lines) + "\n"


def basis_to_dat(basis, densities=False):
    """Text of a basis in the reference's `.dat` format (read_basfile.f90:12-70);
    with `densities` and atomic density blocks in the basis, dmat = 1 and the
    DENSITY section (read_basfile.f90:71-91: per element its label, then each
    row of the block in records of at most 7 values, then ``****``)."""
    els = basis["elements"]
    dens = basis.get("density") if densities else None
    out = ["ELEMENTS %d %d" % (len(els), 1 if dens else 0)]
    for label, shells in els.items():
        out.append("%d" % len(shells))
        out.append("%s 0" % label.capitalize())
        for sh in shells:
            out.append("%s %d 1.00" % (sh["type"].upper(), len(sh["exps"])))
            for k in range(len(sh["exps"])):
                row = [repr(sh["exps"][k]), repr(sh["coefs"][k])]
                if sh["type"] == "l":
                    row.append(repr(sh["coefs_p"][k]))
                out.append("  " + "  ".join(row))
        out.append("****")
    if dens:
        out.append("DENSITY MATRICES")
        for label in els:
            out.append(label.capitalize())
            for row in dens[label]:
                for k in range(0, len(row), 7):
                    out.append(" ".join(repr(float(v)) for v in row[k:k + 7]))
            out.append("****")
    return "\n".join(out) + "\n"


def normalize(zetas, l, coefs):
    """Contraction coefficients including primitive normalisation, normalised
    so that the contracted function has unit self-overlap
    (read_basfile.f90:100-179; for d the x^2 component is the one normalised)."""
    g = []
    for a, c in zip(zetas, coefs):
        if l == 0:
            ofac = (2.0 * a / PI) ** 0.75
        elif l == 1:
            ofac = (2.0 / PI) ** 0.75 * 2.0 * a ** 1.25
        else:
            ofac = (2.0 / PI) ** 0.75 * 4.0 * a ** 1.75 / math.sqrt(3.0)
        g.append(c * ofac)
    nfac = 0.0
    for i, a in enumerate(zetas):
        for j, b in enumerate(zetas):
            p = a + b
            if l == 0:
                nfac += PI32 / math.sqrt(p ** 3) * g[i] * g[j]
            elif l == 1:
                nfac += PI32 / (2.0 * p ** 2 * math.sqrt(p)) * g[i] * g[j]
            else:
                nfac += PI32 * 3.0 / (4.0 * p ** 3 * math.sqrt(p)) * g[i] * g[j]
    s = math.sqrt(nfac)
    return [x / s for x in g]


def atomic_density_guess(mol, basis):
    """Superposition of the basis file's atomic density blocks, the guess
    direct_hartreefock.f90:50-64 builds when USEDENSITYSUPGUESS is set."""
    dens = basis.get("density")
    if not dens:
        raise ValueError("basis %s carries no atomic densities" % basis["name"])
    blocks = [np.array(dens[l]) for l in mol.labels]
    n = sum(b.shape[0] for b in blocks)
    P = np.zeros((n, n))
    k = 0
    for b in blocks:
        m = b.shape[0]
        P[k:k + m, k:k + m] = b
        k += m
    return P


class ShellTable:
    """Flat shell arrays in the layout the C-ABI takes (hfx_mol_desc).

    Arrays: ``shtyp`` (1 s, 2 p, 3 d, 12 L), ``nzet``, ``atsh`` (1-based
    atom), ``aosh`` (1-based first AO), ``xsh`` (nsh, 3) bohr, and
    ``zeta`` / ``djk`` / ``djkl`` (nsh, maxprim).
    """
    MAXPRIM = 10   # molprops.f90:18

    def __init__(self, mol, basis):
        els = basis["elements"]
        shtyp, nzet, atsh, aosh, xsh, zeta, djk, djkl = [], [], [], [], [], [], [], []
        ao = 0
        xb = mol.coords_bohr
        for ia, lab in enumerate(mol.labels):
            if lab not in els:
                raise ValueError("element %s not in basis %s" % (lab, basis["name"]))
            for sh in els[lab]:
                t = SHELL_CODE[sh["type"]]
                n = len(sh["exps"])
                if n > self.MAXPRIM:
                    raise ValueError("more than %d primitives" % self.MAXPRIM)
                z = np.zeros(self.MAXPRIM); c = np.zeros(self.MAXPRIM); cl = np.zeros(self.MAXPRIM)
                z[:n] = sh["exps"]
                if t == 12:
                    cs = normalize(sh["exps"], 0, sh["coefs"])
                    cp = normalize(sh["exps"], 1, sh["coefs_p"])
                    c[:n] = cs
                    cl[:n] = [cp[k] / cs[k] for k in range(n)]
                else:
                    c[:n] = normalize(sh["exps"], {1: 0, 2: 1, 3: 2}[t], sh["coefs"])
                shtyp.append(t); nzet.append(n); atsh.append(ia + 1); aosh.append(ao + 1)
                xsh.append(xb[ia]); zeta.append(z); djk.append(c); djkl.append(cl)
                ao += SHELL_NFUNC[t]
        self.nb = ao
        self.nsh = len(shtyp)
        self.shtyp = np.array(shtyp, dtype=np.int32)
        self.nzet = np.array(nzet, dtype=np.int32)
        self.atsh = np.array(atsh, dtype=np.int32)
        self.aosh = np.array(aosh, dtype=np.int32)
        self.xsh = np.array(xsh, dtype=np.float64).reshape(-1, 3)
        self.zeta = np.array(zeta, dtype=np.float64)
        self.djk = np.array(djk, dtype=np.float64)
        self.djkl = np.array(djkl, dtype=np.float64)
        has = set(self.shtyp.tolist())
        if 2 in has and 3 in has and 12 in has:
            # spread_basis.f90:174-177
            raise ValueError("Mixed P, D and L shells not supported")


# --------------------------------------------------------------------------
# Geometry generators (all coordinates in Angstrom)
# --------------------------------------------------------------------------

def water():
    """RHF water: O at the origin, H at (+-0.7572, 0.5865, 0) Angstrom."""
    return Molecule(["o", "h", "h"], [[0, 0, 0], [0.7572, 0.5865, 0], [-0.7572, 0.5865, 0]], "water")


def benzene(rcc=1.39, rch=1.09):
    """D6h benzene in the xy-plane, carbons first (same layout as the
    reference's benzene.inp:9-20)."""
    labels, xyz = [], []
    for r, lab in ((rcc, "c"), (rcc + rch, "h")):
        for k in range(6):
            a = k * PI / 3.0
            labels.append(lab)
            xyz.append([r * math.cos(a), r * math.sin(a), 0.0])
    return Molecule(labels, xyz, "benzene")


def c60(r66=1.40, r56=1.45):
    """Ideal truncated icosahedron centred at the origin.

    Every atom sits on an icosahedron edge at fraction t from its nearer
    vertex; pentagon edges are t*a, hexagon-hexagon edges (1-2t)*a, so
    t*a = r56 and (1-2t)*a = r66 fix t and the icosahedron edge a.
    """
    a = r56 * 2.0 + r66
    t = r56 / a
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    verts = []
    for s1 in (-1.0, 1.0):
        for s2 in (-1.0, 1.0):
            verts += [(0.0, s1, s2 * phi), (s1, s2 * phi, 0.0), (s2 * phi, 0.0, s1)]
    v = np.array(verts) * (a / 2.0)
    atoms = []
    for i in range(12):
        for j in range(12):
            if i != j and abs(np.linalg.norm(v[i] - v[j]) - a) < 1e-9:
                atoms.append(v[i] + t * (v[j] - v[i]))
    atoms = np.array(atoms)
    assert atoms.shape == (60, 3)
    return Molecule(["c"] * 60, atoms, "c60")


# bond lengths (Angstrom) and the sp3 angle used by the chain generator
BOND = {"cc": 1.54, "ch": 1.09, "co": 1.43, "oh": 0.96, "cn": 1.47, "nh": 1.01}
TETRA = math.acos(-1.0 / 3.0)     # 109.47 degrees
VALENCE = {"c": 4, "n": 3, "o": 2}


def _bond(a, b):
    return BOND["".join(sorted((a, b), key="cnoh".index))]


def _rotate(w, u, ang):
    """w rotated by `ang` about the unit axis u (Rodrigues)."""
    return w * math.cos(ang) + np.cross(u, w) * math.sin(ang) + u * np.dot(u, w) * (1.0 - math.cos(ang))


def chain(nheavy, kind="alkane", jitter=0.0, rng=None):
    """All-trans chain molecule with ideal sp3 geometry, in Angstrom.

    kind: "alkane" (C_n H_2n+2, n = nheavy), "alkanol" (C_n-1 H_2n-1 OH) or
    "amine" (C_n-1 H_2n-1 NH2); the heavy atoms form a planar zigzag in the
    xy-plane, hydrogens complete the tetrahedra.  `jitter` adds seeded
    Gaussian noise (sigma in Angstrom) to every coordinate.
    """
    if nheavy < 1:
        raise ValueError("nheavy >= 1")
    heavy = ["c"] * nheavy
    if kind == "alkanol":
        heavy[-1] = "o"
    elif kind == "amine":
        heavy[-1] = "n"
    elif kind != "alkane":
        raise ValueError("kind must be alkane, alkanol or amine")
    if nheavy == 1 and kind != "alkane":
        raise ValueError("an alkanol or amine needs a carbon")
    half = TETRA / 2.0

    def step(k):   # zigzag bond direction from heavy atom k-1 to k (also for the virtual ends)
        s = 1.0 if k % 2 else -1.0
        return np.array([math.sin(half), s * math.cos(half), 0.0])

    pos = [np.zeros(3)]
    for k in range(1, nheavy):
        pos.append(pos[-1] + _bond(heavy[k - 1], heavy[k]) * step(k))
    labels, xyz = list(heavy), [p.copy() for p in pos]
    for k, el in enumerate(heavy):
        nbrs = []
        if k > 0:
            nbrs.append(-step(k))
        if k + 1 < nheavy:
            nbrs.append(step(k + 1))
        nh = VALENCE[el] - len(nbrs)
        if len(nbrs) == 2:
            b = -(nbrs[0] + nbrs[1]); b /= np.linalg.norm(b)
            n = np.cross(nbrs[0], nbrs[1]); n /= np.linalg.norm(n)
            dirs = [b * math.cos(half) + n * math.sin(half), b * math.cos(half) - n * math.sin(half)]
        elif len(nbrs) == 1:
            u = nbrs[0]
            # first hydrogen where the zigzag would continue (trans), then +-120 deg about the bond
            w0 = -step(k) if k == 0 else step(k + 1)
            dirs = [w0, _rotate(w0, u, 2.0 * PI / 3.0), _rotate(w0, u, -2.0 * PI / 3.0)]
            if el == "n":
                dirs = dirs[1:]
        else:
            t = 1.0 / math.sqrt(3.0)
            dirs = [np.array(v) * t for v in ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))]
        for d in dirs[:nh]:
            labels.append("h")
            xyz.append(pos[k] + _bond(el, "h") * d)
    xyz = np.array(xyz)
    if jitter > 0.0:
        rng = rng if rng is not None else np.random.default_rng(0)
        xyz = xyz + rng.normal(0.0, jitter, xyz.shape)
    name = {"alkane": "C%dH%d" % (nheavy, 2 * nheavy + 2),
            "alkanol": "C%dH%dOH" % (nheavy - 1, 2 * nheavy - 1),
            "amine": "C%dH%dNH2" % (nheavy - 1, 2 * nheavy - 1)}[kind]
    return Molecule(labels, xyz, name)


def batch_molecules(n=512, seed=2024, hmin=3, hmax=20, jitter=0.05):
    """The batched workload: n chain molecules, kinds alkane / alkanol / amine
    in turn, heavy-atom counts uniform in [hmin, hmax], jittered (sigma 0.05 A)."""
    rng = np.random.default_rng(seed)
    kinds = ("alkane", "alkanol", "amine")
    out = []
    for i in range(n):
        nh = int(rng.integers(hmin, hmax + 1))
        out.append(chain(nh, kinds[i % 3], jitter=jitter, rng=rng))
    return out


def serving_requests(rate=20.0, duration=300.0, seed=7, amin=5, amax=120):
    """The serving workload: Poisson arrivals (`rate` per second over
    `duration` s); each request a chain molecule whose atom count is
    log-uniform in [amin, amax], basis STO-3G or 6-31G(d) with equal odds.
    Returns [(arrival_s, Molecule, basis_name)]."""
    rng = np.random.default_rng(seed)
    kinds = ("alkane", "alkanol", "amine")
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        natom = math.exp(rng.uniform(math.log(amin), math.log(amax)))
        kind = kinds[int(rng.integers(0, 3))]
        nheavy = max(1 if kind == "alkane" else 2, int(round((natom - 2.0) / 3.0)))
        basis = "sto-3g" if rng.random() < 0.5 else "6-31g(d)"
        out.append((t, chain(nheavy, kind, jitter=0.05, rng=rng), basis))
