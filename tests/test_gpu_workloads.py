"""The batched and served workloads at their own sizes, and engine
bookkeeping that only a device run shows: C4 (256 columns of 2k-60k material
particles in one engine), C5 (a Poisson stream of Drucker-Prager columns and
von Wolffersdorff embankments of 5k-500k through the request loop), particle
deletion against the oracle, failure flags across layout changes, and the
per-engine interaction variant."""
import ctypes as C

import numpy as np
import pytest

from helpers import rel_err
from loquat_amd import abi, cases
from oracle import lq_oracle as O

FIELDS = ("posd", "pospres", "velrhop", "idep", "str", "stn")


def _same(a: dict, b: dict) -> bool:
    return all(np.array_equal(a[k], b[k], equal_nan=True) for k in FIELDS)


def _alone(case, nsteps):
    from loquat_amd.device import Engine
    e = Engine(case.np, case.derived().ncell, 1)
    rid = e.add(case, nsteps)
    e.step(nsteps)
    out = e.download(rid)
    e.close()
    return out


def _oracle_run(case, nsteps):
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    O.run(p, case.vw, st, nsteps)
    return p, O.unpermute(st)


def _close_to_oracle(case, got, nsteps, tol=1e-4):
    p, ref = _oracle_run(case, nsteps)
    m = slice(p.npb, None)
    x0 = case.posd
    for a, k in enumerate(("x", "y", "z")):
        assert rel_err(got["posd"][m, a] - x0[m, a], getattr(ref, k)[m] - x0[m, a]) < tol, k
    for a, k in enumerate(("vx", "vy", "vz")):
        assert rel_err(got["velrhop"][m, a], getattr(ref, k)[m]) < tol, k
    for c in range(6):
        assert rel_err(got["str"][m, c], ref.s[c, m]) < tol, c


@pytest.mark.gpu
def test_c4_ensemble_batched_equals_alone(gpu):
    """The 256-run ensemble at its full sizes (npm log-uniform 2k-60k, seed
    0), 10 steps in one engine: eight runs -- the smallest, the largest and
    six drawn from a seed -- are bit-identical to the same run alone, and the
    smallest matches the FP64 oracle."""
    from loquat_amd.device import Engine
    nsteps = 10
    runs = cases.ensemble(nruns=256, seed=0, nsteps=nsteps)
    print(f"\nC4: {len(runs)} runs, npm {min(c.npm for c in runs)}..{max(c.npm for c in runs)}, "
          f"sum npm {sum(c.npm for c in runs)}, sum np {sum(c.np for c in runs)}", flush=True)
    e = Engine(sum(c.np for c in runs), sum(c.derived().ncell for c in runs), len(runs))
    ids = [e.add(c) for c in runs]
    e.step(nsteps)
    e.sync()
    assert e.poll() == 0
    npm = np.array([c.npm for c in runs])
    lo, hi = int(npm.argmin()), int(npm.argmax())
    rng = np.random.default_rng(4)
    others = [int(i) for i in rng.choice([i for i in range(len(runs)) if i not in (lo, hi)], 6, replace=False)]
    got = {i: e.download(ids[i]) for i in (lo, hi, *others)}
    for i in (lo, hi, *others):
        assert e.info(ids[i])["steps"] == nsteps
    e.close()
    for i, g in got.items():
        assert _same(g, _alone(runs[i], nsteps)), i
    _close_to_oracle(runs[lo], got[lo], nsteps)


@pytest.mark.gpu
def test_c5_served_mix_equals_alone(gpu):
    """Sixteen requests of the served mix (Poisson arrivals, Drucker-Prager
    columns and von Wolffersdorff embankments 50/50, 5k..500k material
    particles -- the first two at the bounds -- with 12-40 steps each) pass
    through the request loop, arriving a few per tick while others run: every
    result is bit-identical to the same run alone."""
    from loquat_amd.serve import Server
    reqs = cases.served_mix(16, step_scale=0.005, seed=1, extremes=True)
    models = {c.par.constitutive_model for _, c, _ in reqs}
    assert models == {1, 3}
    assert min(c.npm for _, c, _ in reqs) <= 6e3 and max(c.npm for _, c, _ in reqs) >= 4.5e5
    np_cap = sum(c.np for _, c, _ in reqs)
    nc_cap = sum(c.derived().ncell for _, c, _ in reqs)
    s = Server(np_cap, nc_cap, 16, 288 * 10**9, steps_per_tick=5, keep_results=True)
    ids, done, nxt, ticks = [], [], 0, 0
    while len(done) < len(reqs):
        # a few arrivals per tick, as the Poisson times order them
        for _ in range(3):
            if nxt < len(reqs):
                ids.append(s.submit(reqs[nxt][1], reqs[nxt][2]))
                nxt += 1
        done += s.tick()
        ticks += 1
        assert ticks < 500
    assert sorted(done) == sorted(ids)
    for rid, (_, c, n) in zip(ids, reqs):
        assert s.info(rid)["state"] == "done", rid
        got = s.result(rid, c.np)
        assert _same(got, _alone(c, n)), (rid, c.name, c.npm)
    s.close()


@pytest.mark.gpu
def test_compaction_matches_oracle(gpu):
    """Seven material particles start above the grid.  After 4 steps
    lq_engine_compact removes exactly the particles the oracle's cell hash
    puts in cell ncell - 1 (the reference's deletion cell, include/types.h:108);
    after 8 more steps the survivors match the oracle's 12-step run, in which
    the removed particles are never anyone's neighbour."""
    from loquat_amd.device import Engine
    case = cases.perturb(cases.small_column_3d(seed=101), seed=102, v_amp=0.05, s_amp=0.0, rho_amp=0.001)
    p = case.derived()
    gone = p.npb + np.arange(5, 5 + 7 * 37, 37)
    case.posd[gone, 2] = p.domain_size_z + 0.5
    case.pospres[gone, 2] = case.posd[gone, 2].astype(np.float32)
    e = Engine(case.np, p.ncell, 1)
    rid = e.add(case, 12)
    e.step(4)
    assert e.compact() == len(gone)
    e.step(8)
    got = e.download(rid)
    e.close()
    removed = np.nonzero(np.isnan(got["posd"][:, 0]))[0]
    # the oracle's deletion set after the same 4 steps
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    O.run(p, case.vw, st, 4)
    keys = O.cell_keys(p, st)
    want = np.sort(st.orig[(keys == p.ncell - 1) & (np.arange(st.n) >= p.npb)])
    assert np.array_equal(removed, want) and np.array_equal(removed, np.sort(gone))
    O.run(p, case.vw, st, 8, 4)
    ref = O.unpermute(st)
    keep = np.ones(case.np, bool)
    keep[gone] = False
    keep[: p.npb] = False
    x0 = case.posd
    for a, k in enumerate(("x", "y", "z")):
        assert rel_err(got["posd"][keep, a] - x0[keep, a], getattr(ref, k)[keep] - x0[keep, a]) < 1e-4, k
    for a, k in enumerate(("vx", "vy", "vz")):
        assert rel_err(got["velrhop"][keep, a], getattr(ref, k)[keep]) < 1e-4, k
    for c in range(6):
        assert rel_err(got["str"][keep, c], ref.s[c, keep]) < 1e-4, c


@pytest.mark.gpu
def test_failure_flag_survives_layout_changes(gpu):
    """A run that turns non-finite raises its device flag; adding another run
    (which rewrites the run table) before any poll keeps it failed and
    stopped, and the added run steps normally."""
    from loquat_amd.device import Engine
    good = cases.ensemble(nruns=1, seed=8, nsteps=20, npm_lo=800, npm_hi=1500)[0]
    bad = cases.ensemble(nruns=1, seed=9, nsteps=20, npm_lo=800, npm_hi=1500)[0]
    bad.str[bad.par.npb + 10, 0] = np.inf
    e = Engine(good.np + bad.np, good.derived().ncell + bad.derived().ncell, 2)
    rb = e.add(bad, 20)
    e.step(3)
    rg = e.add(good, 5)          # no poll in between
    assert e.failed(rb)
    steps_b = e.info(rb)["steps"]
    assert steps_b <= 3 and e.info(rb)["finished"]
    e.step(5)
    assert e.failed(rb) and e.info(rb)["steps"] == steps_b
    assert not e.failed(rg) and e.info(rg)["steps"] == 5
    assert e.poll() == 1
    got = e.download(rg)
    e.close()
    assert _same(got, _alone(good, 5))


@pytest.mark.gpu
def test_interaction_variant_is_per_engine(gpu):
    """An engine keeps the variant it was created with (the creating thread's
    default) or was given; the thread default only reaches engines created
    later and module-level calls.  Switching an engine's variant mid-run
    drops its captured graphs, so the new variant runs: the result stays
    within FP32 summation order of the single-variant run."""
    from loquat_amd.device import Engine
    lib = abi.lib()
    c = cases.perturb(cases.small_column_3d(seed=41), seed=42)
    try:
        a = Engine(c.np, c.derived().ncell, 1)
        ra = a.add(c, 8)
        assert a.variant() == abi.DEFAULT_INTERACTION_VARIANT
        abi.check(lib.lq_set_interaction_variant(0))
        b = Engine(c.np, c.derived().ncell, 1)
        rb = b.add(c, 8)
        assert b.variant() == 0 and a.variant() == abi.DEFAULT_INTERACTION_VARIANT
        a.step(4)
        a.set_variant(3)
        assert a.variant() == 3 and lib.lq_get_interaction_variant() == 0
        a.step(4)
        b.step(8)
        ga, gb = a.download(ra), b.download(rb)
        a.close()
        b.close()
    finally:
        abi.check(lib.lq_set_interaction_variant(abi.DEFAULT_INTERACTION_VARIANT))
    ref = _alone(c, 8)
    x0 = c.posd
    for g in (ga, gb):
        assert rel_err(g["posd"] - x0, ref["posd"] - x0) < 1e-5
        for k in ("velrhop", "str"):
            assert rel_err(g[k], ref[k]) < 1e-5, k
