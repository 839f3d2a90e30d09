"""Sort interval: a run is sorted every K of its own steps and keeps its cell
table in between, every scan widening its reach by how far the run's
particles have moved since the build.  The neighbour sets must stay those of
a fresh sort at every point of the interval (also for particles crossing
cells, leaving the grid and coming back), the results may differ from
K = 1 only by FP32 summation order, and a run's bits must not depend on when
it joined a batch."""
import numpy as np
import pytest

from helpers import rel_err
from loquat_amd import cases
from oracle import lq_oracle as O

FIELDS = ("posd", "pospres", "velrhop", "idep", "str", "stn")


def _same(a: dict, b: dict) -> bool:
    return all(np.array_equal(a[k], b[k], equal_nan=True) for k in FIELDS)


def _oracle_counts(case, h):
    p = case.derived()
    st = O.OState.from_arrays(h["posd"], h["pospres"], h["velrhop"], h["idep"], h["str"], h["stn"])
    O.initial_sort(p, st)
    g = O.grid_build(p, st)
    _, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    ref = np.zeros_like(nbr)
    ref[st.orig] = nbr
    return ref


def _moving_case(seed: int):
    """A perturbed 3D column with a fast subset: 40 material particles at
    10 m/s in random directions (0.0125 cells per step), 6 starting just
    below the grid's top face moving out, 6 just above it moving in."""
    c = cases.perturb(cases.small_column_3d(seed=seed), seed=seed + 1, v_amp=1.0)
    p = c.derived()
    rng = np.random.default_rng(seed)
    m = p.npb + rng.choice(p.npm, 52, replace=False)
    d = rng.normal(size=(40, 3))
    c.velrhop[m[:40], :3] = (10.0 * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    top = p.cell_num_z * (2.0 / p.i_h)  # the grid's top face (cells start at z = 0)
    c.posd[m[40:46], 2] = top - 0.0005
    c.velrhop[m[40:46], :3] = np.float32([0.0, 0.0, 10.0])
    c.posd[m[46:], 2] = top + 0.0005
    c.velrhop[m[46:], :3] = np.float32([0.0, 0.0, -10.0])
    c.pospres[:, :3] = c.posd.astype(np.float32)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("interval", [4, 16])
def test_kept_table_neighbour_sets_exact(gpu, interval):
    """After every step of an interval and past the next sort, the counts of
    the scan on the kept table equal the oracle's fresh-table counts on the
    engine's own positions, particle by particle."""
    from loquat_amd.device import Engine
    case = _moving_case(201)
    e = Engine(case.np, case.derived().ncell, 1)
    e.set_sort_interval(interval)
    assert e.sort_interval() == interval
    rid = e.add(case, interval + 3)
    npb = case.par.npb
    for k in range(interval + 2):
        got = e.neighbour_counts_stale(rid)
        h = e.download(rid)
        ref = _oracle_counts(case, h)
        assert (got[:npb] == 0).all()
        assert np.array_equal(got[npb:], ref[npb:]), (k, int((got[npb:] != ref[npb:]).sum()))
        e.step(1)
    assert not e.failed(rid)
    e.close()


@pytest.mark.gpu
def test_sort_interval_changes_only_sum_order(gpu):
    """20 steps with a Shepard filter every 7: sorting every step, every 4th
    and every 16th step give the same trajectories to FP32 summation order."""
    from loquat_amd.device import Engine
    case = cases.perturb(cases.small_column_3d(seed=211), seed=212, v_amp=0.5)
    case.par.stress_shepard_filter = 7
    res = {}
    for k in (1, 4, 16):
        e = Engine(case.np, case.derived().ncell, 1)
        e.set_sort_interval(k)
        rid = e.add(case, 20)
        e.step(20)
        res[k] = e.download(rid)
        assert not e.failed(rid)
        e.close()
    x0 = case.posd
    for k in (4, 16):
        assert rel_err(res[1]["posd"] - x0, res[k]["posd"] - x0) < 1e-5, k
        for f in ("velrhop", "str"):
            assert rel_err(res[1][f], res[k][f]) < 1e-5, (k, f)


@pytest.mark.gpu
def test_runs_keep_their_own_sort_schedule(gpu):
    """Three runs join one engine at engine steps 0, 1 and 3 and another
    leaves after step 5 (interval 4, Shepard every 5): each remaining run is
    bit-identical to the same run stepped alone."""
    from loquat_amd.device import Engine
    runs = cases.ensemble(nruns=4, seed=13, nsteps=12, npm_lo=1000, npm_hi=3000)
    for c in runs:
        c.par.stress_shepard_filter = 5
    e = Engine(sum(c.np for c in runs), sum(c.derived().ncell for c in runs), 4)
    ids = [e.add(runs[0], 12), e.add(runs[3], 12)]
    e.step(1)
    ids.insert(1, e.add(runs[1], 12))
    e.step(2)
    ids.insert(2, e.add(runs[2], 12))
    e.step(2)
    e.remove(ids.pop())  # runs[3] leaves
    e.step(12)
    got = [e.download(i) for i in ids]
    e.close()
    for c, g in zip(runs[:3], got):
        a = Engine(c.np, c.derived().ncell, 1)
        r = a.add(c, 12)
        a.step(12)
        alone = a.download(r)
        a.close()
        assert _same(g, alone), c.name


@pytest.mark.gpu
def test_column_order_changes_no_result_bit(gpu):
    """The engine's (x, y) column order (strips of T x-rows) moves particles
    within the sorted arrays only: two runs with grids of 13 and 9 cells
    along x (not multiples of T) give bit-identical states for T = 0, 3, 4
    and 16 over 12 steps (a sort every 4), and the kept-table neighbour counts
    equal the oracle's fresh counts."""
    from loquat_amd.device import Engine
    a = _moving_case(311)
    b = cases.perturb(cases.small_column_3d(nx=10, ny=6, nz=12, seed=313), seed=314)
    res = {}
    for T in (0, 3, 4, 16):
        e = Engine(a.np + b.np, a.derived().ncell + b.derived().ncell, 2)
        e.set_column_order(T)
        assert e.column_order() == T
        e.set_sort_interval(4)
        ra, rb = e.add(a, 12), e.add(b, 12)
        e.step(6)
        if T == 4:
            # mid-interval: the kept table of the tiled order gives the fresh
            # neighbour sets (the next step recomputes the rates it overwrites)
            got = e.neighbour_counts_stale(ra)
            ref = _oracle_counts(a, e.download(ra))
            p = a.derived()
            assert np.array_equal(got[p.npb:], ref[p.npb:])
        e.step(6)
        res[T] = (e.download(ra), e.download(rb))
        e.close()
    for T in (3, 4, 16):
        assert _same(res[0][0], res[T][0]) and _same(res[0][1], res[T][1]), T
