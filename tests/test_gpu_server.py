"""Request loop on the GPU: each result is identical alone and in the mix,
and a diverging run is evicted without disturbing the others."""
import numpy as np
import pytest

from loquat_amd import cases
from loquat_amd.serve import Server


def _alone(case, nsteps):
    from loquat_amd.device import Engine
    e = Engine(case.np, case.derived().ncell, 1)
    rid = e.add(case, nsteps)
    e.step(nsteps)
    out = e.download(rid)
    e.close()
    return out


@pytest.mark.gpu
def test_served_results_equal_alone(gpu):
    runs = cases.ensemble(nruns=6, seed=5, nsteps=40, npm_lo=500, npm_hi=3000)
    steps = [40, 25, 40, 10, 30, 40]
    np_cap = sum(c.np for c in runs)
    nc_cap = sum(c.derived().ncell for c in runs)
    # budget for about three runs at once: the rest queue, join as others leave
    budget = 3 * max(Server.footprint(c) for c in runs)
    s = Server(np_cap, nc_cap, 4, budget, steps_per_tick=5, keep_results=True)
    ids = [s.submit(c, n) for c, n in zip(runs, steps)]
    done = []
    for _ in range(200):
        done += s.tick()
        if len(done) == len(ids):
            break
    assert sorted(done) == sorted(ids)
    for rid, c, n in zip(ids, runs, steps):
        assert s.info(rid)["state"] == "done"
        got = s.result(rid, c.np)
        ref = _alone(c, n)
        for k in ("posd", "velrhop", "str", "stn", "idep"):
            assert np.array_equal(got[k], ref[k]), (rid, k)
    s.close()


@pytest.mark.gpu
def test_diverging_run_evicted(gpu):
    good = cases.ensemble(nruns=1, seed=8, nsteps=20, npm_lo=800, npm_hi=1500)[0]
    bad = cases.ensemble(nruns=1, seed=9, nsteps=20, npm_lo=800, npm_hi=1500)[0]
    bad.str[bad.par.npb + 10, 0] = np.inf
    s = Server(good.np + bad.np, good.derived().ncell + bad.derived().ncell, 2, 10**12, steps_per_tick=5,
               keep_results=True)
    ig, ib = s.submit(good, 20), s.submit(bad, 20)
    done = []
    for _ in range(20):
        done += s.tick()
        if len(done) == 2:
            break
    assert s.info(ib)["state"] == "failed"
    assert s.info(ig)["state"] == "done"
    got = s.result(ig, good.np)
    assert np.array_equal(got["posd"], _alone(good, 20)["posd"])
    s.close()
