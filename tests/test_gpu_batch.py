"""Many-molecule jobs on the GPU.

Shared launches (batch contexts, hfx_batch_*):
* every molecule's builds in a batch context equal its builds alone;
* builds captured in a HIP graph equal uncaptured builds on every replay,
  with eager work and synchronisations between the replays;
* the batched device SCF gives every molecule's standalone energy and
  iteration count, and with mixed precision stays within 1 uEh of FP64;
* the 512-molecule def2-SV(P) workload converges in size-sorted groups, mixed
  precision, each sampled molecule's energy the same in the batch as alone;
* the serving loop (micro-batches) completes a Poisson stream, each sampled
  result the same as the request solved alone.

Per-job contexts (the sequential baseline): concurrent solves equal sequential ones.
"""
import numpy as np
import pytest

from hfcode_amd import batch, fock, molecule as M, scf

pytestmark = pytest.mark.gpu


def _mols():
    return [(M.water(), "sto-3g"),
            (M.chain(3, "alkanol", jitter=0.05, rng=np.random.default_rng(1)), "def2-sv(p)"),
            (M.chain(2, "amine", jitter=0.05, rng=np.random.default_rng(2)), "6-31g(d)"),
            (M.chain(6, "alkane", jitter=0.05, rng=np.random.default_rng(3)), "def2-sv(p)")]


@pytest.fixture(scope="module")
def setup():
    mols = _mols()
    sts = [M.ShellTable(m, M.load_basis(b)) for m, b in mols]
    bb = fock.BatchBuilder(sts)
    S, T, V = bb.onee([m for m, _ in mols])
    alone = []
    for (m, b), st in zip(mols, sts):
        fb = fock.FockBuilder(st)
        s1, t1, v1 = fb.onee(m.znuc.astype(np.float64), m.coords_bohr)
        alone.append(dict(fb=fb, S=s1, H=t1 + v1, P=scf.core_guess(s1, t1 + v1, m.nocc), mol=m))
    yield dict(mols=mols, sts=sts, bb=bb, S=S, H=T + V, alone=alone)
    bb.close()
    for a in alone:
        a["fb"].close()


def test_batch_builds_equal_alone(setup):
    bb, alone = setup["bb"], setup["alone"]
    for k, a in enumerate(alone):
        assert np.abs(bb.unstack(setup["S"], k) - a["S"]).max() < 1e-14
        assert np.abs(bb.unstack(setup["H"], k) - a["H"]).max() < 1e-12
    P = bb.stack([a["P"] for a in alone])
    Z = np.zeros_like(P)
    Fi = bb.initial_diag(Z, P, 5e-7)
    Fn = bb.non_diag(Z, P, 1e-10)
    Fd = bb.diag(Z, P, 1e-10)
    for k, a in enumerate(alone):
        z = np.zeros_like(a["P"])
        fb = a["fb"]
        d = [np.abs(bb.unstack(Fi, k) - fb.initial_diag(z, a["P"], 5e-7)).max(),
             np.abs(bb.unstack(Fn, k) - fb.non_diag(z, a["P"], 1e-10)).max(),
             np.abs(bb.unstack(Fd, k) - fb.diag(z, a["P"], 1e-10)).max()]
        print("batch vs alone, molecule %d (nb %d): %s" % (k, bb.nbs[k], ["%.1e" % x for x in d]), flush=True)
        assert max(d) < 1e-12
        for x, y in zip(bb.schwarz(k), fb.schwarz()):
            assert np.array_equal(x[2], y[2]) and x[3] == y[3]


def test_graph_replays_equal_uncaptured_builds(setup):
    import torch
    bb, alone = setup["bb"], setup["alone"]
    n = bb.nmol
    dev = torch.device("cuda:0")
    P = bb.stack([a["P"] for a in alone])
    bb.initial_diag(np.zeros_like(P), P, 5e-7)
    dP = torch.zeros(P.shape, dtype=torch.float64, device=dev)
    dFg = torch.zeros_like(dP)
    dFe = torch.zeros_like(dP)
    dthr = torch.zeros(3 * n, dtype=torch.float64, device=dev)
    rng = np.random.default_rng(5)
    for rep in range(4):
        dP.copy_(torch.from_numpy(P * (1.0 + 0.1 * rep)))
        t = np.concatenate([rng.choice([1e-10, 1e-9, 5e-7], n), rng.choice([1e-11, 1e-10, 5e-7], n), np.zeros(n)])
        if rep == 3:
            t[n + 1] = np.inf                      # molecule 1: no off-diagonal part
        dthr.copy_(torch.from_numpy(t))
        torch.cuda.synchronize()
        for graph, dF in ((True, dFg), (False, dFe)):
            dF.fill_(7.0)
            bb.build_device(dP.data_ptr(), dF.data_ptr(), dthr.data_ptr(), dthr.data_ptr() + 8 * n, None,
                            zero_f=True, graph=graph)
            torch.cuda.synchronize()
        d = (dFg - dFe).abs().max().item()
        print("replay %d: max|graph - uncaptured| %.2e" % (rep, d), flush=True)
        assert d < 1e-12
        Fe = dFe.cpu().numpy()
        for k, a in enumerate(alone):
            z = np.zeros_like(a["P"])
            pk = P[k, :bb.nbs[k], :bb.nbs[k]] * (1.0 + 0.1 * rep)
            ref = a["fb"].diag(z, pk, t[k])
            if np.isfinite(t[n + k]):
                ref = a["fb"].non_diag(ref, pk, t[n + k])
            assert np.abs(bb.unstack(Fe, k) - ref).max() < 1e-12, (rep, k)


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_scf_equals_alone(setup, graphs):
    bb, alone, mols = setup["bb"], setup["alone"], setup["mols"]
    res = bb.scf(setup["S"], setup["H"], [m.nocc for m, _ in mols], [m.nuclear_repulsion() for m, _ in mols],
                 maxiter=50, diis_start=float("inf"), use_graphs=graphs)
    for k, a in enumerate(alone):
        m = a["mol"]
        dev = a["fb"].scf(a["S"], a["H"], m.nocc, m.nuclear_repulsion(), maxiter=50, diis_start=float("inf"))
        print("molecule %d: batch %.10f (%d it), alone %.10f (%d it)" % (k, res[k].energy, res[k].iterations,
                                                                        dev.energy, dev.iterations), flush=True)
        assert res[k].converged and dev.converged
        assert res[k].iterations == dev.iterations
        assert abs(res[k].energy - dev.energy) < 1e-9


def test_batch_scf_mixed_precision(setup):
    mols, sts = setup["mols"], setup["sts"]
    bb = fock.BatchBuilder(sts)
    try:
        bb.set_mixed(1e-5)
        res = bb.scf(setup["S"], setup["H"], [m.nocc for m, _ in mols], [m.nuclear_repulsion() for m, _ in mols],
                     maxiter=50, diis_start=float("inf"))
    finally:
        bb.close()
    ref = setup["bb"].scf(setup["S"], setup["H"], [m.nocc for m, _ in mols],
                          [m.nuclear_repulsion() for m, _ in mols], maxiter=50, diis_start=float("inf"))
    for r, f in zip(res, ref):
        print("mixed - FP64: %.2e Eh" % (r.energy - f.energy), flush=True)
        assert r.converged and abs(r.energy - f.energy) < 1e-6


def test_batched_512_def2svp_mixed():
    """The 512-molecule workload in groups of 64, mixed precision; every 32nd
    molecule also solved alone (mixed and FP64, batch context of one)."""
    mols = M.batch_molecules(n=512)
    res, wall = batch.run_batched(mols, "def2-sv(p)", group=64, workers=2, split=1e-5)
    conv = [r["converged"] for r in res]
    bad = [(r["name"], r["nb"], r["iterations"], r["status"]) for r in res if not r["converged"]]
    print("512 def2-SV(P) chains: %d converged, %.1f s, %.2f solves/s; not converged: %s" % (
        sum(conv), wall, sum(conv) / wall, bad), flush=True)
    assert sum(conv) >= 510
    for i in range(0, 512, 32):
        alone_mx = batch.solve_group([mols[i]], "def2-sv(p)", split=1e-5)[0]
        alone_64 = batch.solve_group([mols[i]], "def2-sv(p)")[0]
        print("%s nb %d: batch - alone %.2e, mixed - FP64 %.2e" % (
            res[i]["name"], res[i]["nb"], res[i]["energy"] - alone_mx["energy"],
            alone_mx["energy"] - alone_64["energy"]), flush=True)
        if res[i]["converged"]:
            assert alone_mx["converged"] and res[i]["iterations"] == alone_mx["iterations"]
            assert abs(res[i]["energy"] - alone_mx["energy"]) < 1e-8
            assert abs(alone_mx["energy"] - alone_64["energy"]) < 1e-6


def test_serving_micro_batches():
    """20 requests/s for 15 s of the serving mix (5-120 atoms, STO-3G or
    6-31G(d)); every 8th request also solved alone."""
    reqs = M.serving_requests(rate=20.0, duration=15.0, seed=7)
    res, wall = batch.BatchServer(workers=4, group=32).run(reqs)
    assert len(res) == len(reqs)
    assert all(r["error"] is None and r["converged"] for r in res)
    lat = batch.latency_summary(res)
    print("serving %d requests in %.1f s: p50 %.2f s, p99 %.2f s" % (len(res), wall, lat["p50_s"], lat["p99_s"]),
          flush=True)
    for r in res[::8]:
        _, mol, basis = reqs[r["req"]]
        alone = batch.solve_group([mol], basis)[0]
        assert alone["converged"] and r["iterations"] == alone["iterations"]
        assert abs(r["energy"] - alone["energy"]) < 1e-8, r["name"]


def test_per_job_concurrent_solves_match_sequential():
    mols = [M.chain(n, k, jitter=0.05, rng=np.random.default_rng(n))
            for n, k in ((3, "alkane"), (4, "alkanol"), (3, "amine"), (5, "alkane"))]
    seq = [batch.solve(m, "def2-sv(p)") for m in mols]
    conc, wall = batch.run_batch(mols, "def2-sv(p)", workers=4)
    for s, c in zip(seq, conc):
        assert s["converged"] and c["converged"]
        assert abs(s["energy"] - c["energy"]) < 1e-9
