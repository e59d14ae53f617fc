"""Autotuner of the batched workload's launch shapes.

For batch contexts (many small molecules per launch) the best launch width
of each quartet class -- workgroups per compute unit of its kernels -- and,
for the block classes, the number of quartets per chunk of stored integrals
differ from the single-large-molecule defaults, and so does the number of
molecules per batch context.  This tool measures them on the GPU and picks:

* per class, the launch width in PER_CU and then the chunk in CHUNKS with the
  smallest class time (HIP events around the class's launches, summed over
  the workloads and over FP64 and FP32 parts), measured on size-sorted
  groups of the 512-molecule def2-SV(P) workload (small, middle and large
  molecules, mixed precision as that configuration runs) and on the serving
  mix (STO-3G and 6-31G(d) chains);
* the group size in GROUPS with the highest solves/s of batch.run_batched
  on every 4th molecule of the 512.

    python -m hfcode_amd.gen.autotune_batch measure OUT.json   # on the GPU
    python -m hfcode_amd.gen.autotune_batch select OUT.json    # writes batch_launch.json

The table (hfcode_amd/gen/batch_launch.json) is applied to every batch
context by fock.BatchBuilder (hfx_set_class_launch); results do not depend
on it.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "batch_launch.json")
PER_CU = (2, 4, 8, 16)
CHUNKS = (1 << 14, 1 << 16, 1 << 18, 0)       # 0: as many as the integral buffer holds
GROUPS = (16, 32, 64, 128)
DEFAULT = {"per_cu": 0, "chunk": 0}


def load_table(path=TABLE):
    """{"classes": {cls: {"per_cu": n, "chunk": n}}, "group": n} or None."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def _workloads():
    from hfcode_amd import molecule as M
    from hfcode_amd.batch import group_by_size, _nb_of, _basis
    mols = M.batch_molecules(n=512)
    bas = _basis("def2-sv(p)")
    groups = group_by_size([_nb_of(m, bas) for m in mols], 64)
    out = [("def2-sv(p) group %d" % g, [mols[i] for i in groups[g]], "def2-sv(p)", 1e-5) for g in (0, 3, 7)]
    reqs = M.serving_requests(rate=20.0, duration=30.0, seed=11)
    for basis in ("sto-3g", "6-31g(d)"):
        ms = [m for _, m, b in reqs if b == basis][:32]
        out.append(("serving %s" % basis, ms, basis, None))
    return out


def _class_times(bb, P, reps=2):
    """Class times (ms, FP64 + FP32 parts) of one diag + non-diag build."""
    Z = np.zeros_like(P)
    bb.diag(Z, P, 1e-10)
    bb.non_diag(Z, P, 1e-10)            # warm-up: buffers sized
    acc = None
    for _ in range(reps):
        bb.non_diag(Z, P, 1e-10)
        st = bb.stats()
        t = st["class_ms"].sum(axis=1)
        acc = t if acc is None else acc + t
    return acc / reps


def _apply(bb, classes, per_cu=None, chunk=None, table=None):
    from hfcode_amd import fock
    for ci, c in enumerate(fock.CLASSES):
        if table is not None:
            e = table.get(c, DEFAULT)
            bb.set_class_launch(ci, e["per_cu"], e["chunk"])
        else:
            bb.set_class_launch(ci, per_cu or 0, chunk or 0)


def measure(out_path):
    from hfcode_amd import batch, fock, molecule as M, scf
    fock.lib()
    data = {"per_cu": {}, "chunk": {}, "group": {}, "workloads": []}
    for name, mols, basis, split in _workloads():
        sts = [M.ShellTable(m, M.load_basis(basis)) for m in mols]
        bb = fock.BatchBuilder(sts)
        if split:
            bb.set_mixed(split)
        S, T, V = bb.onee(mols)
        P = bb.stack([scf.core_guess(bb.unstack(S, k), bb.unstack(T + V, k), m.nocc) for k, m in enumerate(mols)])
        bb.initial_diag(np.zeros_like(P), P, 5e-7)
        bb.set_profile(True)
        data["workloads"].append(name)
        for pc in PER_CU:
            _apply(bb, fock.CLASSES, per_cu=pc)
            t = _class_times(bb, P)
            for ci, c in enumerate(fock.CLASSES):
                if t[ci] > 0:
                    data["per_cu"].setdefault(c, {}).setdefault(str(pc), 0.0)
                    data["per_cu"][c][str(pc)] += float(t[ci])
            print(name, "per_cu", pc, "total %.2f ms" % t.sum(), flush=True)
        for ch in CHUNKS:
            _apply(bb, fock.CLASSES, chunk=ch)
            t = _class_times(bb, P)
            for ci, c in enumerate(fock.CLASSES):
                if t[ci] > 0:
                    data["chunk"].setdefault(c, {}).setdefault(str(ch), 0.0)
                    data["chunk"][c][str(ch)] += float(t[ci])
            print(name, "chunk", ch, "total %.2f ms" % t.sum(), flush=True)
        bb.close()
    mols = M.batch_molecules(n=512)[::4]
    for g in GROUPS:
        res, wall = batch.run_batched(mols, "def2-sv(p)", group=g, workers=2, split=1e-5, maxiter=64)
        data["group"][str(g)] = sum(r["converged"] for r in res) / wall
        print("group", g, "%.2f solves/s" % data["group"][str(g)], flush=True)
    with open(out_path, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)


def select(data):
    """Table from measurements: per class the fastest launch width, then
    the fastest chunk (block classes only: a class whose times do not depend
    on the chunk keeps 0); the group size with the most solves/s."""
    classes = {}
    for c, times in data["per_cu"].items():
        best = min(times, key=lambda k: (times[k], int(k)))
        classes[c] = {"per_cu": int(best), "chunk": 0}
    split = {}
    flops = os.path.join(HERE, "flops_per_class.json")
    if os.path.exists(flops):
        with open(flops) as fh:
            split = {c: v["split"] for c, v in json.load(fh)["classes"].items()}
    for c, times in data["chunk"].items():
        if c not in classes:
            classes[c] = dict(DEFAULT)
        if not split.get(c, True):      # a fused class stores no integrals: no chunks
            continue
        vals = [times[k] for k in times]
        if max(vals) - min(vals) <= 0.02 * max(vals):      # within noise: no chunk limit
            continue
        classes[c]["chunk"] = int(min(times, key=lambda k: times[k]))
    group = max(data["group"], key=lambda k: (data["group"][k], -int(k))) if data.get("group") else "64"
    return {"about": "launch shapes of batch contexts picked by hfcode_amd/gen/autotune_batch.py on the "
                     "512-molecule def2-SV(P) workload and the serving mix",
            "classes": classes, "group": int(group)}


def main(argv):
    if len(argv) != 2 or argv[0] not in ("measure", "select"):
        print(__doc__)
        return 2
    if argv[0] == "measure":
        measure(argv[1])
    else:
        with open(argv[1]) as fh:
            tab = select(json.load(fh))
        with open(TABLE, "w") as fh:
            json.dump(tab, fh, indent=1, sort_keys=True)
            fh.write("\n")
        print("wrote", TABLE)
    return 0


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    sys.exit(main(sys.argv[1:]))
