"""Many-molecule jobs on one GPU: batched SCF solves and a serving loop.

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
