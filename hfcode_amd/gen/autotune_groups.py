"""Autotuner of the block-class kernels' grouping of L-part combinations.

The block classes emit one kernel per group of part combinations
(gen_eri.Gen.group_plan): more combinations per kernel mean fewer passes
over the primitive pairs and Boys evaluations, but more live registers.
This tool measures, per class, which accumulator budget is fastest, for the
FP64 kernels and for the FP32 ones separately, on the C60 / 6-31G(d) Fock
build, and writes the choice to group_budgets.json, which the generator
reads.  Three steps:

    python -m hfcode_amd.gen.autotune_groups build            # CPU: library variants
    python -m hfcode_amd.gen.autotune_groups measure OUT.json # GPU: time every class
    python -m hfcode_amd.gen.autotune_groups select OUT.json  # write group_budgets.json

The variants are built under hfcode_amd/_tune/<name>/ (not part of the
product); the product library is rebuilt afterwards by hfcode_amd.build.
"""
import ctypes
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TUNE = os.path.join(ROOT, "hfcode_amd", "_tune")
# variant name -> (FP64 budget, FP32 budget), contracted accumulators per kernel
VARIANTS = {"b16": (16, 32), "b32": (32, 64), "b64": (64, 128), "b128": (128, 256), "b256": (256, 512)}


def variant_lib(name):
    return os.path.join(TUNE, name, "libhfx.so")


def build_variants(names=None):
    sys.path.insert(0, ROOT)
    from hfcode_amd import build
    for name in names or VARIANTS:
        d = os.path.join(TUNE, name)
        build.build(gen=os.path.join(d, "gen"), obj_dir=os.path.join(d, "obj"), lib=variant_lib(name),
                    budgets=VARIANTS[name])


def _measure_one(lib_path, steps=3):
    """Per-class kernel times (ms per build) of the FP64 build and of the FP32
    part of the mixed build (split 1e-5) of C60 / 6-31G(d) at its converged
    density, with the given library."""
    sys.path.insert(0, ROOT)
    import numpy as np
    import torch
    from hfcode_amd import fock, molecule as M, scf
    torch.cuda.init()
    fock._lib = fock.configure(ctypes.CDLL(lib_path))
    mol = M.c60()
    bas = M.load_basis("6-31g(d)")
    fb = fock.FockBuilder(M.ShellTable(mol, bas))
    S, T, V = fb.onee(mol.znuc.astype(np.float64), mol.coords_bohr)
    H = T + V
    P = scf.rhf(fb, S, H, mol.nocc, mol.nuclear_repulsion(), P0=M.atomic_density_guess(mol, bas), maxiter=40).P
    dev = torch.device("cuda:0")
    dP = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    dF = torch.empty_like(dP)
    fb.set_profile(True)
    fb.set_overlap(False)       # lists one after another: a class's span is its own kernels only
    out = {}
    for key, split in (("fp64", 0.0), ("mixed", 1e-5)):
        fb.set_mixed(split)
        acc = None
        for it in range(steps + 1):
            dF.zero_()
            fb.diag_device(dP.data_ptr(), dF.data_ptr(), 1e-10)
            fb.non_diag_device(dP.data_ptr(), dF.data_ptr(), 1e-10)
            cm = fb.stats()["class_ms"]
            if it:                      # the first build warms up
                acc = cm if acc is None else acc + cm
        acc = acc / steps
        out[key] = {c: [float(acc[i, 0]), float(acc[i, 1])] for i, c in enumerate(fock.CLASSES)}
    fb.close()
    return out


def measure(out_path, names=None):
    res = {}
    for name in names or VARIANTS:
        t0 = time.time()
        # one process per variant: each loads its own copy of the library
        r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "_one", variant_lib(name)],
                           stdout=subprocess.PIPE, text=True, check=True)
        res[name] = json.loads(r.stdout.strip().splitlines()[-1])
        print("measured %s in %.0f s" % (name, time.time() - t0), file=sys.stderr, flush=True)
    with open(out_path, "w") as fh:
        json.dump({"variants": VARIANTS, "workload": "C60 6-31G(d) Fock build, converged density, thr 1e-10",
                   "class_ms": res}, fh, indent=1)


def select(measure_path, out_path=None):
    sys.path.insert(0, ROOT)
    from hfcode_amd.gen import gen_eri
    m = json.load(open(measure_path))
    variants, cms = m["variants"], m["class_ms"]
    budgets, report = {}, {}
    for cls in gen_eri.CLASSES:
        if gen_eri.Gen(cls).fused_cost() <= gen_eri.SPLIT_COST:
            continue
        t64 = {v: cms[v]["fp64"][cls][0] for v in cms}
        t32 = {v: cms[v]["mixed"][cls][1] for v in cms}
        if max(t64.values()) <= 0.0:
            continue
        v64 = min(t64, key=t64.get)
        v32 = min(t32, key=t32.get) if max(t32.values()) > 0.0 else v64   # no FP32 quartets measured
        budgets[cls] = [variants[v64][0], variants[v32][1]]
        report[cls] = {"fp64_ms": t64, "fp32_ms": t32}
    out_path = out_path or gen_eri.GROUP_BUDGETS_FILE
    with open(out_path, "w") as fh:
        json.dump({"workload": m["workload"], "budgets": budgets, "measured_ms": report}, fh, indent=1,
                  sort_keys=True)
    print("wrote", out_path)


if __name__ == "__main__":
    cmd = sys.argv[1] if len(sys.argv) > 1 else ""
    if cmd == "build":
        build_variants(sys.argv[2:] or None)
    elif cmd == "measure":
        measure(sys.argv[2], sys.argv[3:] or None)
    elif cmd == "select":
        select(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else None)
    elif cmd == "_one":
        print(json.dumps(_measure_one(sys.argv[2])), flush=True)
    else:
        sys.exit(__doc__)
