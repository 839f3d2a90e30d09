"""Request loop on the CPU fake backend: admission by HBM footprint, strict
FIFO order, continuous batching, eviction of failed runs."""
import numpy as np
import pytest

from loquat_amd import abi, cases
from loquat_amd.serve import Server


def small(nx, seed=0):
    return cases.small_column_3d(nx=nx, ny=4, nz=6, seed=seed)


def footprint(case):
    p = case.derived()
    return abi.lib().lq_engine_bytes_per_run(p.np, p.ncell)


def test_admission_by_footprint_and_fifo():
    a, b, c = small(6, 1), small(8, 2), small(5, 3)
    fa, fb, fc = footprint(a), footprint(b), footprint(c)
    # budget holds a + c but not a + b: b waits for a, and c (behind b) waits too
    budget = fa + max(fb, fc) - 1
    assert fa + fc <= budget < fa + fb
    s = Server(10**6, 10**7, 8, budget, steps_per_tick=10, fake=True)
    ia, ib, ic = s.submit(a, 30), s.submit(b, 40), s.submit(c, 20)
    done = s.tick()                      # admits a only; b does not fit
    assert done == []
    assert s.info(ia)["state"] == "running"
    assert s.info(ib)["state"] == "queued" and s.info(ic)["state"] == "queued"
    assert s.stats()["bytes"] == fa
    s.tick()
    done = s.tick()                      # a finishes after 30 steps
    assert done == [ia]
    assert s.info(ia)["state"] == "done" and s.info(ia)["steps"] == 30
    assert s.stats()["bytes"] == 0
    s.tick()                             # b then c admitted (FIFO), both fit together
    assert s.info(ib)["state"] == "running" and s.info(ic)["state"] == "running"
    order = []
    for _ in range(5):
        order += s.tick()
    assert order == [ic, ib]             # c needs 20 steps, b 40
    assert s.stats() == dict(queued=0, running=0, bytes=0, npm=0)


def test_rejects_requests_larger_than_budget():
    a = small(6)
    s = Server(10**6, 10**7, 4, footprint(a) - 1, steps_per_tick=5, fake=True)
    rid = s.submit(a, 10)
    assert s.info(rid)["state"] == "rejected"
    assert s.tick() == []


def test_failed_run_is_evicted_without_disturbing_others():
    good, bad = small(6, 1), small(6, 2)
    bad.velrhop[bad.par.npb + 3, 0] = np.nan
    s = Server(10**6, 10**7, 4, 10**12, steps_per_tick=5, fake=True)
    ig, ib = s.submit(good, 15), s.submit(bad, 15)
    first = s.tick()
    assert first == [ib] and s.info(ib)["state"] == "failed"
    assert s.info(ig)["state"] == "running"
    rest = s.tick() + s.tick()
    assert rest == [ig] and s.info(ig)["state"] == "done"


def test_run_slots_limit_concurrency():
    runs = [small(5, i) for i in range(5)]
    s = Server(10**6, 10**7, 2, 10**12, steps_per_tick=10, fake=True)
    ids = [s.submit(r, 10) for r in runs]
    finished = []
    for _ in range(3):
        assert s.stats()["running"] <= 2
        finished += s.tick()
    assert finished == ids[:6][: len(finished)]
    while len(finished) < 5:
        finished += s.tick()
    assert finished == ids


def test_invalid_config():
    with pytest.raises(abi.LqError):
        Server(0, 10, 1, 10, 1, fake=True)


@pytest.mark.parametrize("defect", ["model", "vw", "wall"])
def test_invalid_request_is_rejected_and_does_not_block_the_queue(defect):
    """A request that lq_engine_add_run would refuse (constitutive model out
    of range, model 3 without its constants, a wall particle outside the
    grid) is rejected at submission; the requests queued behind it run."""
    a, bad, c = small(6, 1), small(6, 2), small(5, 3)
    if defect == "model":
        bad.par.constitutive_model = 7
    elif defect == "vw":
        bad.par.constitutive_model = 3
        bad.vw = abi.LqVw()
    else:
        bad.posd[0, 0] = -1.0
        bad.pospres[0, 0] = -1.0
    s = Server(10**6, 10**7, 4, 10**12, steps_per_tick=10, fake=True)
    ia, ib, ic = s.submit(a, 20), s.submit(bad, 20), s.submit(c, 20)
    assert s.info(ib)["state"] == "rejected"
    done = []
    for _ in range(3):
        done += s.tick()
    assert sorted(done) == sorted([ia, ic])
    assert s.info(ia)["state"] == "done" and s.info(ic)["state"] == "done"
