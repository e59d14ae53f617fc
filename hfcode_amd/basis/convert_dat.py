"""Convert a basis file in the reference's `.dat` format to this package's JSON.

The `.dat` layout is the one `read_basfile` parses (read_basfile.f90:12-70):
``ELEMENTS n dmat`` then, per element, the shell count, the element label,
and per shell ``T nprim scale`` followed by ``nprim`` lines of exponent and
coefficient (plus a p coefficient for ``L`` shells), closed by ``****``.
With ``dmat=1`` the file ends with one atomic density block per element
(read_basfile.f90:71-91, used by the superposition-of-atomic-densities
guess, direct_hartreefock.f90:50-64); they are kept as ``density``.

The JSON keeps the raw (unnormalised) published coefficients; normalisation is
done by ``hfcode_amd.molecule.normalize`` exactly as read_basfile.f90:100-179
does it.  Usage::

    python -m hfcode_amd.basis.convert_dat SRC.dat OUT.json NAME
"""
import json
import sys


def parse_dat(text):
    toks = [ln.split() for ln in text.splitlines()]
    toks = [t for t in toks if t]
    pos = 0
    head = toks[pos]; pos += 1
    nel = int(head[1])
    elements = {}
    for _ in range(nel):
        nshell = int(toks[pos][0]); pos += 1
        label = toks[pos][0].lower(); pos += 1
        shells = []
        for _ in range(nshell):
            typ = toks[pos][0].lower(); nprim = int(toks[pos][1]); pos += 1
            exps, cs, cps = [], [], []
            for _ in range(nprim):
                row = toks[pos]; pos += 1
                exps.append(float(row[0].replace('D', 'E').replace('d', 'e')))
                cs.append(float(row[1].replace('D', 'E').replace('d', 'e')))
                if typ == 'l':
                    cps.append(float(row[2].replace('D', 'E').replace('d', 'e')))
            sh = {"type": typ, "exps": exps, "coefs": cs}
            if typ == 'l':
                sh["coefs_p"] = cps
            shells.append(sh)
        assert toks[pos][0].startswith('****'), toks[pos]
        pos += 1
        elements[label] = shells
    densities = {}
    if int(head[2]) == 1 and pos < len(toks):
        assert toks[pos][0].upper() == "DENSITY", toks[pos]
        pos += 1
        ncomp = {"s": 1, "p": 3, "d": 5, "l": 4}
        for label in elements:
            lab = toks[pos][0].lower(); pos += 1
            assert lab == label, (lab, label)
            n = sum(ncomp[sh["type"]] for sh in elements[label])
            vals = []
            while len(vals) < n * n:
                vals += [float(v) for v in toks[pos]]
                pos += 1
            assert toks[pos][0].startswith('****'), toks[pos]
            pos += 1
            densities[label] = [vals[i * n:(i + 1) * n] for i in range(n)]
    return elements, densities


def main(argv):
    src, out, name = argv[1], argv[2], argv[3]
    with open(src) as fh:
        elements, densities = parse_dat(fh.read())
    doc = {"name": name, "elements": elements}
    if densities:
        doc["density"] = densities
    with open(out, 'w') as fh:
        json.dump(doc, fh, indent=1)


if __name__ == '__main__':
    main(sys.argv)
