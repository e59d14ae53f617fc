"""Many-molecule jobs on one GPU: batched SCF solves and a serving loop.

Shared launches (the default): molecules are grouped by basis size; each
group is one batch context (``fock.BatchBuilder``: every quartet class of
every molecule of the group in one launch of its kernels) and one batched
SCF on the GPU (``BatchBuilder.scf``: all of the group's S, H, P, F, DIIS
history and orbitals resident in device memory, each iteration's build and
post-build sequence replayed from HIP graphs).  A few groups run at once on
their own streams.  ``solve_group`` / ``run_batched`` / ``BatchServer``.

Per job (kept as the sequential baseline and for comparison):

Each job is a full RHF solve (the driver of direct_hartreefock.f90:98-223:
``hfcode_amd.scf.rhf`` on the host around the GPU builds, or with
``device_scf`` all of it on the GPU through ``FockBuilder.scf``) with its own
Fock-build context (``FockBuilder``: device pair data, Schwarz tables, work
lists and one HIP stream).  A pool of host worker threads runs the jobs; the
library calls release the interpreter lock, so the jobs' Fock builds run
concurrently on their own streams and fill the GPU that one small molecule
alone would leave mostly idle.  Host linear algebra runs with one BLAS
thread per worker.

Admission: a job is started only while the device memory its context would
take (work-list capacity, integral buffers; estimated from its shell pairs
before any allocation) fits in the budget next to the running jobs.
"""
import contextlib
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor, as_completed

import numpy as np

from . import fock, molecule as M, scf

_basis_cache = {}
_basis_lock = threading.Lock()


def _basis(name):
    with _basis_lock:
        if name not in _basis_cache:
            _basis_cache[name] = M.load_basis(name)
        return _basis_cache[name]


def memory_estimate(st):
    """Device bytes a context for shell table `st` may take: the quartet
    lists (capacity = candidate pair-pairs of every class, two lists with
    mixed precision off: one), the integral buffer of the block classes and
    its partial maxima (bounded by the library at 5.2 GB + 16 doubles per
    quartet of a chunk while several contexts are live; a context that grew
    its buffer while it was the only one may hold up to 5.2 GB more, which
    the budget's distance to the device's memory covers), and nb^2
    matrices."""
    types = {}
    for i in range(st.nsh):
        for j in range(i + 1):
            a, b = sorted((int(st.shtyp[i]), int(st.shtyp[j])), key=[1, 2, 3, 12].index)
            types[(a, b)] = types.get((a, b), 0) + 1
    npairs = list(types.values())
    cap = 0
    for x in range(len(npairs)):
        for y in range(x, len(npairs)):
            cap += npairs[x] * npairs[y] if x != y else npairs[x] * (npairs[x] - 1) // 2
    eblk = min((cap + 255) // 256 * 256 * 625, (1 << 20) * 625)      # doubles
    vmax = min(1 << 22, (eblk // 16 + 255) // 256 * 256) * 16         # doubles
    return 8 * cap + 8 * (eblk + vmax) + 8 * 8 * st.nb * st.nb + (1 << 27)


DIIS_START = float("inf")   # DIIS from the second iteration on


def solve(mol, basis_name, maxiter=50, thr=scf.INTTHRESH, diis_start=DIIS_START, split=None,
          device_scf=False):
    """One RHF solve on the GPU; returns a result dict.

    `split`: quartets whose Schwarz x density bound is below it run in FP32
    (FockBuilder.set_mixed); None = all FP64.  `device_scf`: the whole SCF
    iteration on the GPU (FockBuilder.scf, hfx_scf_run); the default runs the
    host driver (scf.rhf) around the GPU builds, which is faster when many
    small jobs share the GPU (measured: 512 def2-SV(P) chains 3.61 solves/s
    host driver, 2.99 device driver; rocSOLVER's dsyevd at nb ~200 is many
    small launches, which compete with the other jobs' builds).

    Guess: the superposition of the basis file's atomic densities where the
    basis has them (STO-3G, 6-31G(d): the reference's USEDENSITYSUPGUESS,
    direct_hartreefock.f90:50-64), else the core Hamiltonian (def2-SV(P),
    whose basis file has none).  From the core guess, the reference's
    schedule -- plain iterations until max|FPS - SPF| < 0.3, then DIIS
    (direct_hartreefock.f90:145-150) -- oscillates for many chain molecules,
    with or without a level shift (methanol 6-31G(d): E stays near
    -91.5 Eh; with the same Fock builds the converged energy is -115.03 Eh);
    the same DIIS (8 vectors, 1.05 damping, 1.5 cap) started at the second
    iteration converges them, so the jobs run it that way."""
    t0 = time.perf_counter()
    basis = _basis(basis_name)
    st = M.ShellTable(mol, basis)
    fb = fock.FockBuilder(st)
    try:
        if split:
            fb.set_mixed(split)
        S, T, V = fb.onee(mol.znuc.astype(np.float64), mol.coords_bohr)
        H = T + V
        P0 = M.atomic_density_guess(mol, basis) if basis.get("density") else None
        t1 = time.perf_counter()
        if device_scf:
            res = fb.scf(S, H, mol.nocc, mol.nuclear_repulsion(), maxiter=maxiter, intthresh=thr,
                         diis_start=diis_start, P0=P0)
        else:
            res = scf.rhf(fb, S, H, mol.nocc, mol.nuclear_repulsion(), maxiter=maxiter, intthresh=thr,
                          diis_start=diis_start, P0=P0)
        t2 = time.perf_counter()
    finally:
        fb.close()
    t3 = time.perf_counter()
    last = res.history[-1] if res.history else (0, None, None, None)
    return dict(name=mol.title, natom=mol.natom, nb=st.nb, basis=basis_name, energy=res.energy,
                iterations=res.iterations, converged=bool(res.converged), setup_s=t1 - t0, scf_s=t2 - t1,
                close_s=t3 - t2, fock_s=res.t_fock, split=split, last_dE=last[2], last_dP=last[3])


def _blas_one_thread():
    """Context: one BLAS thread per calling thread (threadpoolctl), if available."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:     # BLAS keeps its own setting
        return contextlib.nullcontext()
    return threadpool_limits(1)


def run_batch(mols, basis_name, workers=16, maxiter=50, progress=None, split=None, device_scf=False):
    """Solve every molecule of `mols`; returns (results in input order, wall
    seconds).  `progress(done, total, elapsed_s)` is called as jobs finish."""
    fock.lib()                      # load the library (and initialise the runtime) once
    with _blas_one_thread():
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=workers) as ex:
            futs = [ex.submit(solve, m, basis_name, maxiter, split=split, device_scf=device_scf) for m in mols]
            if progress is not None:
                done = 0
                for _ in as_completed(futs):
                    done += 1
                    progress(done, len(futs), time.perf_counter() - t0)
            out = [f.result() for f in futs]
        wall = time.perf_counter() - t0
    return out, wall


class Server:
    """Serving loop: requests (arrival time, molecule, basis) are admitted in
    arrival order when a worker is free and the memory budget allows; the
    latency of a request is completion time - arrival time."""

    def __init__(self, workers=16, mem_budget=200 * (1 << 30), maxiter=50, split=None, device_scf=False):
        self.workers = workers
        self.mem_budget = mem_budget
        self.maxiter = maxiter
        self.split = split
        self.device_scf = device_scf
        self._lock = threading.Condition()
        self._inflight_mem = 0
        self._inflight = 0
        self._progress = None
        self._total = 0

    def _run_one(self, req, t_start, need, out):
        arrival, mol, basis = req
        try:
            r = solve(mol, basis, self.maxiter, split=self.split, device_scf=self.device_scf)
            r["error"] = None
        except Exception as e:          # a failed job is reported, the loop goes on
            r = dict(name=mol.title, natom=mol.natom, basis=basis, error=repr(e), converged=False)
        done = time.perf_counter() - t_start
        r.update(arrival_s=arrival, done_s=done, latency_s=done - arrival)
        with self._lock:
            self._inflight_mem -= need
            self._inflight -= 1
            out.append(r)
            if self._progress is not None:
                self._progress(len(out), self._total, done)
            self._lock.notify_all()

    def run(self, requests, progress=None):
        """Replay `requests` in real time; returns (results, wall seconds).
        `progress(done, total, elapsed_s)` is called as requests finish."""
        fock.lib()
        out = []
        threads = []
        self._progress, self._total = progress, len(requests)
        with _blas_one_thread():
            t_start = time.perf_counter()
            for req in requests:
                arrival, mol, basis = req
                now = time.perf_counter() - t_start
                if arrival > now:
                    time.sleep(arrival - now)
                need = memory_estimate(M.ShellTable(mol, _basis(basis)))
                with self._lock:
                    while self._inflight >= self.workers or (
                            self._inflight > 0 and self._inflight_mem + need > self.mem_budget):
                        self._lock.wait()
                    self._inflight += 1
                    self._inflight_mem += need
                th = threading.Thread(target=self._run_one, args=(req, t_start, need, out), daemon=True)
                th.start()
                threads.append(th)
            for th in threads:
                th.join()
            wall = time.perf_counter() - t_start
        out.sort(key=lambda r: r["arrival_s"])
        return out, wall


def run_sequential(jobs, maxiter=50, split=None):
    """Solve (molecule, basis) jobs one at a time on one thread (the plain
    one-job-at-a-time GPU path); returns (results, wall seconds)."""
    fock.lib()
    out = []
    t0 = time.perf_counter()
    for mol, basis in jobs:
        out.append(solve(mol, basis, maxiter, split=split))
    return out, time.perf_counter() - t0


def latency_summary(results):
    lat = np.array([r["latency_s"] for r in results])
    return dict(n=len(lat), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), max_s=float(lat.max()))


# ---- shared launches: groups of molecules in batch contexts -------------------

def _nb_of(mol, basis):
    return M.ShellTable(mol, basis).nb


def solve_group(mols, basis_name, maxiter=50, diis_start=DIIS_START, split=None, use_graphs=True, on_done=None):
    """RHF of every molecule of `mols` (one basis) in one batch context and
    one batched device SCF; returns result dicts in input order.
    `on_done(i, result)` is called as molecule i finishes (converged or not).
    Guess and DIIS schedule as :func:`solve`; `split` as FockBuilder.set_mixed
    (full builds only; incremental builds stay FP64)."""
    t0 = time.perf_counter()
    basis = _basis(basis_name)
    sts = [M.ShellTable(m, basis) for m in mols]
    bb = fock.BatchBuilder(sts)
    out = [None] * len(mols)
    try:
        if split:
            bb.set_mixed(split)
        S, T, V = bb.onee(mols)
        H = T + V
        P0 = bb.stack([M.atomic_density_guess(m, basis) for m in mols]) if basis.get("density") else None
        t1 = time.perf_counter()

        def done(i, r):
            out[i] = dict(name=mols[i].title, natom=mols[i].natom, nb=sts[i].nb, basis=basis_name, energy=r.energy,
                          iterations=r.iterations, converged=r.converged, status=r.status, split=split,
                          setup_s=t1 - t0, scf_s=time.perf_counter() - t1, group=len(mols))
            if on_done is not None:
                on_done(i, out[i])
        bb.scf(S, H, [m.nocc for m in mols], [m.nuclear_repulsion() for m in mols], P0=P0, maxiter=maxiter,
               diis_start=diis_start, use_graphs=use_graphs, on_done=done)
    finally:
        bb.close()
    return out


def group_by_size(nbs, group):
    """Index groups of at most `group` molecules, consecutive in basis size
    (so that a group's stacked matrices carry little padding)."""
    order = sorted(range(len(nbs)), key=lambda i: nbs[i])
    return [order[k:k + group] for k in range(0, len(order), group)]


def run_batched(mols, basis_name, group=None, workers=2, maxiter=100, split=None, use_graphs=True, progress=None):
    """Solve every molecule of `mols` in size-sorted groups of `group`
    molecules (default: the tuned group size of batch_launch.json, else 64),
    `workers` groups at once; returns (results in input order, wall seconds).
    At most 100 iterations: from the core guess the reference's DIIS schedule
    takes 50 iterations for some long chains even with the reference's own
    Fock builds, and one of the 512 def2-SV(P) chains of the batched workload
    (a C19H40, nb 346) needs 64 in FP64."""
    fock.lib()
    if group is None:
        from .gen import autotune_batch
        tab = autotune_batch.load_table()
        group = int(tab["group"]) if tab else 64
    basis = _basis(basis_name)
    nbs = [_nb_of(m, basis) for m in mols]
    out = [None] * len(mols)
    lock = threading.Lock()
    count = [0]
    t0 = time.perf_counter()

    def one(idx):
        def done(i, r):
            with lock:
                out[idx[i]] = r
                count[0] += 1
                if progress is not None:
                    progress(count[0], len(mols), time.perf_counter() - t0)
        solve_group([mols[i] for i in idx], basis_name, maxiter=maxiter, split=split, use_graphs=use_graphs,
                    on_done=done)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(one, g) for g in group_by_size(nbs, group)]:
            f.result()
    return out, time.perf_counter() - t0


def group_memory_estimate(sts):
    """Device bytes a batch context and its batched SCF may take for shell
    tables `sts`: the quartet lists and integral buffer of every molecule
    (memory_estimate) and the SCF's stacked matrices ((15 + 2 ndiis) per
    molecule at the group's largest nb)."""
    nb = max(st.nb for st in sts)
    return sum(memory_estimate(st) for st in sts) + 8 * 31 * nb * nb * len(sts)


class BatchServer:
    """Serving loop with continuous batching.  Requests (arrival time,
    molecule, basis) queue as they arrive; whenever one of `workers` slots is
    free, the oldest queued request and the queued ones of its basis closest
    to it in size (basis sizes within `size_ratio`, at most `group`, within
    the memory budget next to the running groups) start as one micro-batch: one batch context and one batched SCF, so their builds share
    every launch.  Micro-batches overlap on the GPU on their own streams; a
    request's result is reported as soon as its own SCF finishes, while the
    rest of its micro-batch goes on.  Latency = completion - arrival."""

    def __init__(self, workers=4, group=32, mem_budget=200 * (1 << 30), maxiter=50, split=None, use_graphs=True,
                 size_ratio=2.0):
        self.workers = workers
        self.group = group
        self.size_ratio = size_ratio
        self.mem_budget = mem_budget
        self.maxiter = maxiter
        self.split = split
        self.use_graphs = use_graphs

    def run(self, requests, progress=None):
        """Replay `requests` in real time; returns (results sorted by arrival, wall seconds)."""
        fock.lib()
        cond = threading.Condition()
        queue = []                  # (index, arrival, molecule, basis, shell table)
        out = []
        state = dict(running=0, mem=0, arrived=0)
        t_start = time.perf_counter()

        def worker_run(batch, need):
            idx = [b[0] for b in batch]
            mols = [b[2] for b in batch]
            basis = batch[0][3]

            def done(i, r):
                t = time.perf_counter() - t_start
                r.update(req=idx[i], arrival_s=batch[i][1], done_s=t, latency_s=t - batch[i][1], error=None)
                with cond:
                    out.append(r)
                    if progress is not None:
                        progress(len(out), len(requests), t)
            try:
                solve_group(mols, basis, maxiter=self.maxiter, split=self.split, use_graphs=self.use_graphs,
                            on_done=done)
            except Exception as e:   # a failed micro-batch is reported, the loop goes on
                t = time.perf_counter() - t_start
                with cond:
                    for k in range(len(batch)):
                        if not any(o.get("req") == idx[k] for o in out):
                            out.append(dict(req=idx[k], name=mols[k].title, natom=mols[k].natom, basis=basis, error=repr(e),
                                            converged=False, arrival_s=batch[k][1], done_s=t,
                                            latency_s=t - batch[k][1]))
            with cond:
                state["running"] -= 1
                state["mem"] -= need
                cond.notify_all()

        def dispatcher():
            threads = []
            while True:
                with cond:
                    while True:
                        if not queue and state["arrived"] == len(requests):
                            for th in threads:
                                th.join(0)
                            return threads
                        if queue and state["running"] < self.workers:
                            # the oldest request and the queued ones closest to it in
                            # size (same basis, basis size within `size_ratio`): a
                            # micro-batch steps in lockstep, so a small molecule
                            # should not share it with a much larger one
                            basis, nb0 = queue[0][3], queue[0][4].nb
                            cands = [q for q in queue if q[3] == basis and
                                     max(q[4].nb, nb0) <= self.size_ratio * min(q[4].nb, nb0)]
                            cands = cands[:1] + sorted(cands[1:], key=lambda q: abs(math.log(q[4].nb / nb0)))
                            batch, need = [], 0
                            for q in cands:
                                if len(batch) >= self.group:
                                    break
                                est = group_memory_estimate([b[4] for b in batch] + [q[4]])
                                if batch and state["mem"] + est > self.mem_budget:
                                    break
                                batch.append(q)
                                need = est
                            if batch and (state["running"] == 0 or state["mem"] + need <= self.mem_budget):
                                for q in batch:
                                    queue.remove(q)
                                state["running"] += 1
                                state["mem"] += need
                                break
                        cond.wait(0.05)
                th = threading.Thread(target=worker_run, args=(batch, need), daemon=True)
                th.start()
                threads.append(th)

        disp_threads = []
        disp = threading.Thread(target=lambda: disp_threads.extend(dispatcher()), daemon=True)
        with _blas_one_thread():
            disp.start()
            for i, (arrival, mol, basis) in enumerate(requests):
                now = time.perf_counter() - t_start
                if arrival > now:
                    time.sleep(arrival - now)
                st = M.ShellTable(mol, _basis(basis))
                with cond:
                    queue.append((i, arrival, mol, basis, st))
                    state["arrived"] += 1
                    cond.notify_all()
            disp.join()
            for th in disp_threads:
                th.join()
            wall = time.perf_counter() - t_start
        out.sort(key=lambda r: r["arrival_s"])
        return out, wall
