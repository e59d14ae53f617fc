"""Many-molecule jobs on the GPU: concurrent solves (one context and stream
each) give the energies of the same solves run one after another, and the
chain-molecule generator's molecules converge."""
import numpy as np
import pytest

from hfcode_amd import batch, molecule as M


@pytest.mark.gpu
def test_concurrent_solves_match_sequential():
    mols = [M.chain(n, k, jitter=0.05, rng=np.random.default_rng(n))
            for n, k in ((3, "alkane"), (4, "alkanol"), (3, "amine"), (5, "alkane"), (2, "alkanol"), (6, "amine"))]
    seq = [batch.solve(m, "def2-sv(p)") for m in mols]
    conc, wall = batch.run_batch(mols, "def2-sv(p)", workers=6)
    for s, c in zip(seq, conc):
        print(s["name"], s["nb"], s["energy"], c["energy"], s["iterations"], flush=True)
        assert s["converged"] and c["converged"]
        assert abs(s["energy"] - c["energy"]) < 1e-9


@pytest.mark.gpu
def test_serving_loop_completes_every_request():
    reqs = M.serving_requests(rate=20.0, duration=1.0, seed=3, amax=30)
    res, wall = batch.Server(workers=4).run(reqs)
    assert len(res) == len(reqs)
    assert all(r["error"] is None and r["converged"] for r in res)
    lat = batch.latency_summary(res)
    assert lat["p99_s"] >= lat["p50_s"] > 0
