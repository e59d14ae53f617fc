#!/usr/bin/env python3
"""Fock builds/s for RHF C60 / 6-31G(d) on one MI355X.

Workload: the ideal truncated icosahedron (6-6 bonds 1.40 A, 5-6 bonds
1.45 A), 6-31G(d) with spherical d as the reference uses it (840 basis
functions, 240 shells, 28,920 shell pairs), a fixed density P already in
HBM.  One step = one full two-electron Fock build F = H + G(P):
fock_build_diag + fock_build_non_diag at thr = 1e-10 (the reference's
intthresh), all 21 quartet classes of 6-31G(d).

Timed twice: plain FP64, and mixed precision (quartets whose
Schwarz x density bound is below 1e-5 in FP32, accumulation FP64); the
energy error of the mixed build is E(P, F_mixed) - E(P, F_fp64).
`value` is the mixed builds/s when that error is <= 1 uEh, else FP64.

The roofline object is for the class kernel with the largest device time
(HIP events on the library's stream, averaged over the timed steps): its
algorithmic FLOPs (operation counts of the generated kernel code x the
primitive / contracted quartets it evaluated) over its duration, against the
FP64 vector peak of MI355X.  cpu_baseline times the reference Fortran
(oracle/_ref/libhfref.so, built from the reference's sources) on one C60
build on this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Two more workloads print their own JSON line:

    python bench.py --workload batch [--batch-n 512] [--workers 16]
        RHF solves/s over n all-trans chain molecules (alkanes, alkanols,
        amines, 3-20 heavy atoms, jittered), def2-SV(P), core guess, each a
        full SCF on the GPU Fock build; many molecules in flight at once.
    python bench.py --workload serve [--rate 20] [--duration 60] [--workers 16]
        serving loop: Poisson arrivals of chain molecules (5-120 atoms,
        log-uniform; STO-3G or 6-31G(d)), p50/p99 latency and solves/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector, AMD public spec
PEAK_HBM_TBS = 8.0          # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c60_631gd_hbm_traffic.json")
PEAK_FP32_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


_OUT = None


def emit(out):
    """The one JSON result line, on the original stdout."""
    print(json.dumps(out), file=_OUT or sys.stdout, flush=True)


def _claim_stdout():
    """Keep the process's stdout for the result line alone: native libraries
    (the reference Fortran of the CPU baseline prints its thread count when
    it exits) write to file descriptor 1, which from here on is stderr."""
    global _OUT
    sys.stdout.flush()
    _OUT = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)


def _class_kernels(cls):
    """Names of the kernels a class's FP64 span covers."""
    path = os.path.join(ROOT, "hfcode_amd", "gen", "flops_per_class.json")
    split = False
    if os.path.exists(path):
        with open(path) as fh:
            split = json.load(fh)["classes"][cls]["split"]
    if split:
        # the digestion form the library picks for a one-molecule FP64 list (table of
        # measured classes, with long enough bra runs, which C60's lists have)
        from hfcode_amd.gen.gen_eri import DIG1_USE
        dig = ("dig1_%s<double>" % cls) if DIG1_USE.get(cls, 0) & 1 else ("dig_%s<false,double>" % cls)
        return "class (%s|%s) FP64 span: eri_%s_c*<double> + %s" % (cls[:2], cls[2:], cls, dig)
    return "class (%s|%s) FP64: eri_%s<double,false>" % (cls[:2], cls[2:], cls)


def setup(density, thr):
    from hfcode_amd import fock, molecule as M, scf
    mol = M.c60()
    bas = M.load_basis("6-31g(d)")
    st = M.ShellTable(mol, bas)
    fb = fock.FockBuilder(st)
    S, T, V = fb.onee(mol.znuc.astype(np.float64), mol.coords_bohr)
    H = T + V
    unuc = mol.nuclear_repulsion()
    t0 = time.time()
    if density == "scf":
        # superposition-of-atomic-densities guess, as the reference does with
        # USEDENSITYSUPGUESS (direct_hartreefock.f90:50-64)
        P0 = M.atomic_density_guess(mol, bas)
        res = scf.rhf(fb, S, H, mol.nocc, unuc, P0=P0, maxiter=40)
        P = res.P
        info = dict(scf_energy=res.energy, scf_iterations=res.iterations, scf_converged=res.converged)
    else:
        from hfcode_amd.scf import core_guess
        P = core_guess(S, H, mol.nocc)
        fb.initial_diag(np.zeros_like(H), P, 5e-7)
        info = {}
    info["setup_s"] = time.time() - t0
    return mol, st, fb, S, H, P, unuc, info


def time_builds(fb, dP, dH, dF, thr, steps, warmup, torch):
    for _ in range(warmup):
        dF.copy_(dH)
        fb.diag_device(dP.data_ptr(), dF.data_ptr(), thr)
        fb.non_diag_device(dP.data_ptr(), dF.data_ptr(), thr)
    torch.cuda.synchronize()
    class_ms = None
    class_flops = None
    t0 = time.perf_counter()
    for _ in range(steps):
        dF.copy_(dH)
        fb.diag_device(dP.data_ptr(), dF.data_ptr(), thr)
        fb.non_diag_device(dP.data_ptr(), dF.data_ptr(), thr)
        st = fb.stats()          # waits for the build; events of the non-diagonal kernels
        class_ms = st["class_ms"] if class_ms is None else class_ms + st["class_ms"]
        class_flops = st["class_flops"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, class_ms / steps, class_flops, st


def host_cpu():
    """(threads to use, description) of the host's CPU share: OMP_NUM_THREADS
    when the environment sets it (a GPU box's share of its host), else every
    CPU the process may run on; the model name from /proc/cpuinfo."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else avail
    threads = min(threads, avail)
    model, phys = None, set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if ":" not in line:
                    if "core id" in cur:
                        phys.add((cur.get("physical id"), cur["core id"]))
                    cur = {}
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
                if k == "model name" and model is None:
                    model = v
    except OSError:
        pass
    return threads, dict(model=model, cpus_available=avail, physical_cores=len(phys) or None,
                         omp_num_threads_env=env or None)


def cpu_reference(mol, bas, P, thr, threads):
    """One C60 build by the reference Fortran on the host (wall clock)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hfref
    if not hfref.available():
        return None, None
    r = hfref.RefMolecule.from_molecule(mol, bas)
    z = np.zeros_like(P)
    r.initial_diag(z, P, 5e-7)
    t0 = time.perf_counter()
    F = r.diag(z, P, thr)
    F = r.non_diag(F, P, thr)
    dt = time.perf_counter() - t0
    r.free()
    return dt, F


def _device_used_bytes():
    """Device-wide bytes in use (hipMemGetInfo through torch), or None."""
    try:
        import torch
        free, total = torch.cuda.mem_get_info(0)
        return total - free
    except Exception:
        return None


def _energy_check(seq, mine):
    """Max |E(job) - E(same job solved alone in FP64)| over the jobs both
    lists hold (matched by position) and how many of them converged."""
    d = [abs(a["energy"] - b["energy"]) for a, b in zip(seq, mine)
         if a.get("converged") and b.get("converged")]
    return (max(d) if d else None), len(d)


def _sequential(jobs):
    """One-at-a-time plain FP64 GPU solves: the per-job path (host SCF driver
    around a single-molecule context) and a batch context of one (device
    driver); the faster of the two is the baseline."""
    from hfcode_amd import batch
    seq_host, wall_host = batch.run_sequential(jobs)
    t0 = time.perf_counter()
    seq_dev = [batch.solve_group([m], b)[0] for m, b in jobs]
    wall_dev = time.perf_counter() - t0
    if wall_dev < wall_host:
        return seq_dev, wall_dev, "batch context of one (device SCF driver, graphs)", dict(per_job_host=wall_host)
    return seq_host, wall_host, "single-molecule context, host SCF driver", dict(batch_of_one=wall_dev)


def run_batch(a):
    from hfcode_amd import batch, molecule as M
    mols = M.batch_molecules(n=a.batch_n)
    split = a.split if a.split is not None else 1e-5      # the configuration names mixed precision
    sample = list(range(0, len(mols), max(1, len(mols) // a.seq_n)))[:a.seq_n]
    log("batch: sequential FP64 solves of %d sampled molecules" % len(sample))
    seq, seq_wall, seq_kind, seq_other = _sequential([(mols[i], "def2-sv(p)") for i in sample])
    log("batch: sequential %.1f s (%s)" % (seq_wall, seq_kind))
    peak = [0]

    def progress(done, total, el):
        u = _device_used_bytes()
        if u:
            peak[0] = max(peak[0], u)
        if done % 64 == 0 or done == total:
            log("batch: %d/%d solved, %.1f s" % (done, total, el))
    maxiter = 100
    if a.path == "batched":
        res, wall = batch.run_batched(mols, "def2-sv(p)", group=a.group, workers=a.workers, split=split or None,
                                      progress=progress, maxiter=maxiter)
    else:
        res, wall = batch.run_batch(mols, "def2-sv(p)", workers=a.workers, progress=progress, split=split or None,
                                    device_scf=a.scf == "device", maxiter=maxiter)
    nbs = [r["nb"] for r in res]
    its = [r["iterations"] for r in res]
    conv = sum(r["converged"] for r in res)
    value = conv / wall          # converged solves only
    unconverged = [dict(name=r["name"], nb=r["nb"], iterations=r["iterations"], last_dE=r.get("last_dE"),
                        last_dP=r.get("last_dP")) for r in res if not r["converged"]]
    seq_rate = len(sample) / seq_wall
    dE, ndE = _energy_check(seq, [res[i] for i in sample])
    out = {
        "metric": "RHF solves/s, %d chain molecules def2-SV(P)" % len(mols),
        "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * wall, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if split else "f64",
        "data": "synthetic: all-trans alkanes / alkanols / amines, 3-20 heavy atoms, "
                "ideal bonds + 0.05 A Gaussian jitter, seed 2024",
        "config": {"workload": "batched RHF solves (core guess, DIIS from iteration 2, thr 1e-10, <= %d iterations)" % maxiter,
                   "basis": "def2-SV(P)", "molecules": len(mols), "nb_min": min(nbs), "nb_max": max(nbs),
                   "nb_mean": round(float(np.mean(nbs)), 1), "path": a.path, "group": a.group,
                   "workers": a.workers, "mixed_split": split, "parallelism": "single GPU"},
        "converged": conv, "unconverged": unconverged,
        "iterations_mean": round(float(np.mean(its)), 2), "iterations_max": max(its),
        "sequential_fp64_solves_per_s": round(seq_rate, 4),
        "sequential_kind": seq_kind, "sequential_other_wall_s": seq_other,
        "sequential_sample": "every %dth molecule (%d), one at a time, FP64" % (sample[1] - sample[0] if len(sample) > 1 else 1, len(sample)),
        "vs_sequential": round(value / seq_rate, 3),
        "energy_error_max_Eh": dE, "energy_error_jobs": ndE,
        "energy_error_note": "batch result (mixed precision) minus the same molecule's sequential FP64 solve",
        "peak_device_bytes": peak[0] or None,
        "wall_s": round(wall, 3),
    }
    emit(out)


def run_serve(a):
    from hfcode_amd import batch, molecule as M
    reqs = M.serving_requests(rate=a.rate, duration=a.duration)
    split = a.split if a.split else None
    nseq = min(a.seq_n, len(reqs))
    log("serve: sequential FP64 solves of the first %d requests" % nseq)
    seq, seq_wall, seq_kind, seq_other = _sequential([(m, b) for _, m, b in reqs[:nseq]])
    log("serve: sequential %.1f s (%s)" % (seq_wall, seq_kind))
    if a.path == "batched":
        srv = batch.BatchServer(workers=a.workers, group=a.group, split=split)
    else:
        srv = batch.Server(workers=a.workers, split=split, device_scf=a.scf == "device")
    peak = [0]

    def progress(done, total, el):
        u = _device_used_bytes()
        if u:
            peak[0] = max(peak[0], u)
        if done % 100 == 0 or done == total:
            log("serve: %d/%d requests done, %.1f s" % (done, total, el))
    res, wall = srv.run(reqs, progress=progress)
    lat = batch.latency_summary(res)
    ok = [r for r in res if not r.get("error") and r.get("converged")]
    rate = len(ok) / wall
    seq_rate = nseq / seq_wall
    dE, ndE = _energy_check(seq, res[:nseq])
    out = {
        "metric": "serving RHF p99 latency, Poisson %.0f/s chain molecules STO-3G/6-31G(d)" % a.rate,
        "value": round(lat["p99_s"], 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * wall, 1), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if split else "f64",
        "data": "synthetic: Poisson arrivals, seed 7; chain molecules 5-120 atoms log-uniform",
        "config": {"workload": "serving loop, full RHF per request (atomic-density guess, thr 1e-10)",
                   "rate_per_s": a.rate, "duration_s": a.duration, "path": a.path, "group": a.group,
                   "workers": a.workers, "mixed_split": split, "parallelism": "single GPU"},
        "requests": len(reqs), "completed": len(ok), "converged": sum(r.get("converged", False) for r in res),
        "p50_s": round(lat["p50_s"], 4), "p99_s": round(lat["p99_s"], 4), "mean_s": round(lat["mean_s"], 4),
        "max_s": round(lat["max_s"], 4), "solves_per_s": round(rate, 3),
        "sequential_fp64_solves_per_s": round(seq_rate, 4),
        "sequential_kind": seq_kind, "sequential_other_wall_s": seq_other,
        "sequential_sample": "first %d requests, one at a time, FP64" % nseq,
        "vs_sequential": round(rate / seq_rate, 3),
        "energy_error_max_Eh": dE, "energy_error_jobs": ndE,
        "peak_device_bytes": peak[0] or None,
        "wall_s": round(wall, 2),
    }
    emit(out)


def main():
    _claim_stdout()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--thr", type=float, default=1e-10)
    ap.add_argument("--split", type=float, default=None,
                    help="mixed-precision split (default 1e-5 for c60, 0 = FP64 for batch and serve)")
    ap.add_argument("--density", choices=["scf", "core"], default="scf")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="threads of the CPU baseline (default: OMP_NUM_THREADS, else every available CPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["c60", "batch", "serve"], default="c60")
    ap.add_argument("--chunk", type=int, default=0, help="quartets per block-class chunk (0: library default)")
    ap.add_argument("--batch-n", type=int, default=512)
    ap.add_argument("--workers", type=int, default=None,
                    help="batch/serve: groups (batched path) or jobs (per-job path) in flight")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--seq-n", type=int, default=32, help="batch/serve: jobs of the sequential FP64 baseline")
    ap.add_argument("--scf", choices=["host", "device"], default="host",
                    help="batch/serve per-job path: SCF driver of each job (host: scf.rhf, device: hfx_scf_run)")
    ap.add_argument("--path", choices=["batched", "per-job"], default="batched",
                    help="batch/serve: shared launches in batch contexts, or one context per job")
    ap.add_argument("--group", type=int, default=64, help="batch/serve: molecules per batch context, at most")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the build would shard its screened quartet lists by cost and all-reduce X once per build)",
              file=sys.stderr)
        sys.exit(2)
    if a.workers is None:
        a.workers = (2 if a.workload == "batch" else 4) if a.path == "batched" else 16
    if a.workload == "serve" and a.group == 64 and a.path == "batched":
        a.group = 32
    if a.workload == "batch":
        return run_batch(a)
    if a.workload == "c60" and a.split is None:
        a.split = 1e-5
    if a.workload == "serve":
        return run_serve(a)
    import torch
    from hfcode_amd import fock, molecule as M
    torch.cuda.init()
    mol, st, fb, S, H, P, unuc, info = setup(a.density, a.thr)
    log("setup done", info)
    dev = torch.device("cuda:0")
    dP = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    dH = torch.from_numpy(np.ascontiguousarray(H)).to(dev)
    dF = torch.empty_like(dH)
    fb.set_profile(True)
    if a.chunk:
        fb.set_chunk(a.chunk)
    # FP64
    fb.set_mixed(0.0)
    dt64, cms64, cfl64, st64 = time_builds(fb, dP, dH, dF, a.thr, a.steps, a.warmup, torch)
    F64 = dF.cpu().numpy()
    E64 = 0.5 * float(np.sum(P * (H + F64))) + unuc
    # mixed
    fb.set_mixed(a.split)
    dtmx, cmsmx, cflmx, stmx = time_builds(fb, dP, dH, dF, a.thr, a.steps, a.warmup, torch)
    Fmx = dF.cpu().numpy()
    Emx = 0.5 * float(np.sum(P * (H + Fmx))) + unuc
    fb.set_mixed(0.0)
    # one FP64 build at the tighter cut-off the reference configuration also uses
    dt12, _, _, st12 = time_builds(fb, dP, dH, dF, 1e-12, 1, 1, torch)
    # class spans with the block-class lists one after another (untimed): in
    # the timed builds a class's first and last kernels overlap its
    # neighbours' digestion and integrals, which its span then includes
    fb.set_overlap(False)
    _, cms_alone, _, _ = time_builds(fb, dP, dH, dF, a.thr, 1, 1, torch)
    fb.set_overlap(True)
    err = abs(Emx - E64)
    b64 = a.steps / dt64
    bmx = a.steps / dtmx
    use_mixed = err <= 1e-6
    value = bmx if use_mixed else b64
    # roofline: the quartet class with the largest device time in the FP64
    # build.  A block class is several kernels (its combination kernels and
    # its digestion, chunk by chunk), timed as one span with HIP events on the
    # library's stream around all of them; its algorithmic FLOPs are the
    # frozen per-class counts (hfcode_amd/gen/flops_per_class.json) times the
    # quartets the screening passed.
    k = int(np.argmax(cms64[:, 0]))
    ach = cfl64[k, 0] / (cms64[k, 0] * 1e-3) / 1e12
    ach_alone = cfl64[k, 0] / (cms_alone[k, 0] * 1e-3) / 1e12
    cname = fock.CLASSES[k]
    kinfo = _class_kernels(cname)
    tot_fl = st64["flops_fp64"]
    q64 = int(st64["quartets"][:, 0].sum())
    pq64 = int(st64["prim_quartets"][:, 0].sum())
    out = {
        "metric": "Fock builds/s C60/6-31G(d) mixed vs FP64 @<=1uEh",
        "value": round(value, 4),
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 / value, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if use_mixed else "f64",
        "data": "synthetic: ideal C60 geometry, %s density" % ("converged FP64 SCF" if a.density == "scf" else
                                                                "core-Hamiltonian guess"),
        "config": {"workload": "RHF C60 6-31G(d) Fock build (diag + non-diag), 5d spherical, nb=%d, thr=%g"
                   % (st.nb, a.thr), "molecule": "C60", "basis": "6-31G(d)", "nb": st.nb, "thr": a.thr,
                   "mixed_split": a.split, "parallelism": "single GPU"},
        "fp64_builds_per_s": round(b64, 4),
        "mixed_builds_per_s": round(bmx, 4),
        "mixed_speedup": round(bmx / b64, 3),
        "mixed_energy_error_Eh": err,
        "fp64_energy_Eh": E64,
        "quartets_per_build_fp64": q64,
        "prim_quartets_per_build_fp64": pq64,
        "quartets_per_s_fp64": q64 * b64,
        "build_tflops_fp64": tot_fl * b64 / 1e12,
        "roofline": {
            "bound": "valu-fp64",
            "kernel": kinfo,
            "class": cname,
            "achieved": round(ach, 3),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(ach / PEAK_FP64_TFLOPS, 4),
            "traffic": None,
            "flops_per_launch": cfl64[k, 0],
            "avg_launch_ms": round(float(cms64[k, 0]), 4),
            "span_note": "the class's span in the timed builds, first to last of its launches; its first and "
                         "last kernels run beside the neighbouring lists' kernels",
            "alone": {"avg_launch_ms": round(float(cms_alone[k, 0]), 4), "achieved": round(ach_alone, 3),
                      "frac": round(ach_alone / PEAK_FP64_TFLOPS, 4),
                      "note": "one untimed FP64 build with the block-class lists one after another"},
        },
        "class_ms_fp64": {fock.CLASSES[i]: round(float(cms64[i, 0]), 3) for i in range(46) if cms64[i, 0] > 0},
        "class_ms_mixed": {fock.CLASSES[i]: [round(float(cmsmx[i, 0]), 3), round(float(cmsmx[i, 1]), 3)]
                           for i in range(46) if cmsmx[i, 0] + cmsmx[i, 1] > 0},
        "fp32_quartet_fraction": round(float(stmx["quartets"][:, 1].sum()) / max(1.0, float(stmx["quartets"].sum())), 4),
        "fp32_prim_quartet_fraction": round(float(stmx["prim_quartets"][:, 1].sum())
                                            / max(1.0, float(stmx["prim_quartets"].sum())), 4),
        "fp64_builds_per_s_thr1e-12": round(1.0 / dt12, 4),
        "quartets_per_build_thr1e-12": int(st12["quartets"][:, 0].sum()),
    }
    # HBM bytes of the roofline kernel: FETCH_SIZE + WRITE_SIZE counter runs of this
    # workload (rocprofv3 cannot run inside the timed process), when they match it
    if os.path.exists(TRAFFIC_FILE) and a.density == "scf" and a.thr == 1e-10 and not a.chunk:
        with open(TRAFFIC_FILE) as fh:
            tr = json.load(fh)
        if tr.get("class") == cname:
            t_s = out["roofline"]["avg_launch_ms"] * 1e-3
            out["roofline"]["traffic"] = tr["traffic_bytes"]
            out["roofline"]["traffic_source"] = tr["source"]
            out["roofline"]["hbm_TBps"] = round(tr["traffic_bytes"] / t_s / 1e12, 3)
            out["roofline"]["hbm_frac"] = round(tr["traffic_bytes"] / t_s / 1e12 / PEAK_HBM_TBS, 4)
    out.update({k_: v for k_, v in info.items()})
    if not a.no_cpu:
        bas = M.load_basis("6-31g(d)")
        threads, cpu = host_cpu()
        if a.cpu_threads:
            threads = a.cpu_threads
        dt, Fr = cpu_reference(mol, bas, P, a.thr, threads)
        if dt is not None:
            out["cpu_baseline"] = {"value": round(1.0 / dt, 5), "unit": "Fock builds/s", "cores": threads,
                                   "kind": "reference",
                                   "sample": "one full C60 build (fock_build_diag + fock_build_non_diag, thr %g) "
                                             "of the reference Fortran, flang -O3 -fopenmp, OpenMP threads = the "
                                             "process's CPU share, wall clock" % a.thr,
                                   "cpu_model": cpu["model"], "cpus_available": cpu["cpus_available"],
                                   "host_physical_cores": cpu["physical_cores"],
                                   "omp_num_threads_env": cpu["omp_num_threads_env"]}
            out["gpu_fp64_vs_cpu"] = round(b64 * dt, 2)
            out["fock_maxabs_vs_reference"] = float(np.abs((Fr + H) - F64).max())
    emit(out)


if __name__ == "__main__":
    main()
