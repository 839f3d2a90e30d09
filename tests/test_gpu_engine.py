"""Engine (time loop) on the GPU: trajectory against the FP64 oracle, and
batching invariants -- a run's result is bit-identical alone, in a batch,
with graphs or direct launches, and across joins/leaves of other runs."""
import os

import numpy as np
import pytest

from helpers import rel_err
from loquat_amd import abi, cases
from oracle import lq_oracle as O


def _engine(cases_, steps=None, graphs=True, spare=()):
    """Engine holding cases_ (and room for the runs in spare)."""
    from loquat_amd.device import Engine
    np_cap = sum(c.np for c in (*cases_, *spare))
    nc = sum(c.derived().ncell for c in (*cases_, *spare))
    e = Engine(np_cap, nc, max_runs=max(1, len(cases_) + len(spare)))
    e.set_graphs(graphs)
    ids = [e.add(c, steps) for c in cases_]
    return e, ids


def _same(a: dict, b: dict) -> bool:
    return all(np.array_equal(a[k], b[k], equal_nan=True) for k in ("posd", "pospres", "velrhop", "idep", "str", "stn"))


def gentle_case(model: int, seed: int = 21):
    """Geostatic column with a small random velocity field: a state in which
    all three constitutive models stay stable (random stress perturbations
    put surface particles into tension, where hypoplastic stiffness vanishes
    and both implementations diverge chaotically)."""
    from test_gpu_parity import WU
    c = cases.perturb(cases.small_column_3d(nx=14, ny=8, nz=20, seed=seed), seed=seed + 1, v_amp=0.05, s_amp=0.0,
                      rho_amp=0.001)
    if model == 2:
        c.par.constitutive_model = 2
        (c.par.model_par_1, c.par.model_par_2, c.par.model_par_3, c.par.model_par_4, c.par.model_par_5) = WU
    elif model == 3:
        from loquat_amd import abi
        c.par.constitutive_model = 3
        c.vw = abi.LqVw(**cases.VW_SAND)
        rng = np.random.default_rng(seed)
        c.idep[c.par.npb:, 2] = (0.7 * (1 + rng.uniform(-0.05, 0.05, c.npm))).astype(np.float32)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("model", [1, 2, 3])
def test_engine_vs_oracle_steps(gpu, model):
    """20 full steps (with Shepard every 5) from the same start."""
    case = gentle_case(model)
    case.par.stress_shepard_filter = 5
    nsteps = 20
    e, (rid,) = _engine([case], nsteps)
    e.step(nsteps)
    got = e.download(rid)
    assert e.info(rid)["steps"] == nsteps and e.info(rid)["finished"]
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    O.run(p, case.vw, st, nsteps)
    ref = O.unpermute(st)
    npb = p.npb
    m = slice(npb, None)
    x0 = case.posd
    for a, k in enumerate(("x", "y", "z")):
        assert rel_err(got["posd"][m, a] - x0[m, a], getattr(ref, k)[m] - x0[m, a]) < 1e-4, k
    for a, k in enumerate(("vx", "vy", "vz")):
        assert rel_err(got["velrhop"][m, a], getattr(ref, k)[m]) < 1e-4, k
    assert rel_err(got["velrhop"][m, 3] - p.rho0, ref.rho[m] - p.rho0) < 1e-3
    for c in range(6):
        assert rel_err(got["str"][m, c], ref.s[c, m]) < 1e-4, c
    e.close()


@pytest.mark.gpu
def test_engine_artificial_pressure_vs_oracle(gpu):
    """20 steps with artificial-pressure = 1, beside a run without it in
    the same batch: a cohesive Drucker-Prager column whose top layers start
    in isotropic tension.  The batch integrates in separate kernels (runs
    with artificial pressure keep them), the run without the term alone
    inside its interaction launches: the two give the same bits."""
    case = gentle_case(1, seed=31)
    case.par.model_par_5 = 4.0e3          # DP-kcohesion: tension is admissible
    top = case.posd[:, 2] > np.quantile(case.posd[case.par.npb:, 2], 0.8)
    top[:case.par.npb] = False
    case.str[top, :3] += 1.5e3
    case.pospres[:, 3] = -case.str[:, :3].sum(1) / 3.0
    case.par.artificial_pressure = 1
    case.par.artificial_pressure_coefficient = 0.3
    other = gentle_case(1, seed=32)
    nsteps = 20
    e, (rid, oid) = _engine([case, other], nsteps)
    e.step(nsteps)
    got = e.download(rid)
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    O.run(p, case.vw, st, nsteps)
    ref = O.unpermute(st)
    m = slice(p.npb, None)
    x0 = case.posd
    for a, k in enumerate(("x", "y", "z")):
        assert rel_err(got["posd"][m, a] - x0[m, a], getattr(ref, k)[m] - x0[m, a]) < 1e-4, k
    for a, k in enumerate(("vx", "vy", "vz")):
        assert rel_err(got["velrhop"][m, a], getattr(ref, k)[m]) < 1e-4, k
    for c in range(6):
        assert rel_err(got["str"][m, c], ref.s[c, m]) < 1e-4, c
    # the run without the term is unaffected by its batch neighbour
    e2, (rid2,) = _engine([other], nsteps)
    e2.step(nsteps)
    assert _same(e.download(oid), e2.download(rid2))
    e.close()
    e2.close()


@pytest.mark.gpu
def test_batch_equals_alone(gpu):
    runs = cases.ensemble(nruns=5, seed=3, nsteps=30, npm_lo=500, npm_hi=4000)
    alone = []
    for c in runs:
        e, (rid,) = _engine([c])
        e.step(30)
        alone.append(e.download(rid))
        e.close()
    extra = cases.ensemble(nruns=1, seed=99, nsteps=30, npm_lo=500, npm_hi=4000)[0]
    e, ids = _engine(runs, spare=(extra,))
    e.step(12)
    # a run leaves and another joins mid-way; the others must not notice
    e.remove(ids[1])
    e.add(extra)
    e.step(18)
    for i, rid in enumerate(ids):
        if i == 1:
            continue
        assert _same(e.download(rid), alone[i]), i
    e.close()


@pytest.mark.gpu
def test_graphs_equal_direct(gpu):
    c = cases.perturb(cases.small_column_3d(seed=31, shepard=3), seed=32)
    out = []
    for graphs in (True, False):
        e, (rid,) = _engine([c], 10, graphs=graphs)
        e.step(10)
        out.append(e.download(rid))
        e.close()
    assert _same(out[0], out[1])


@pytest.mark.gpu
def test_interaction_variants_agree_in_engine(gpu):
    from loquat_amd import abi
    c = cases.perturb(cases.small_column_3d(seed=41), seed=42)
    res = []
    vs = (0, *range(3, abi.INTERACTION_VARIANTS))
    for v in vs:
        abi.check(abi.lib().lq_set_interaction_variant(v))
        e, (rid,) = _engine([c], 5)
        e.step(5)
        res.append(e.download(rid))
        e.close()
    abi.check(abi.lib().lq_set_interaction_variant(abi.DEFAULT_INTERACTION_VARIANT))
    for r in res[1:]:
        for k in ("posd", "velrhop", "str"):
            assert rel_err(res[0][k], r[k]) < 1e-5
    # variants 3 to 11 run the same pair code on the same pairs in the same
    # order; 12 visits the stencil columns in another order, 13 to 15 and 17
    # to 19 in a third (partial flushes, staging and masking keep the order);
    # 16 splits the sums over two lanes
    by_v = dict(zip(vs, res))
    for v in range(4, 12):
        assert _same(by_v[3], by_v[v]), v
    for v in (14, 15, 17, 18, 19):
        assert _same(by_v[13], by_v[v]), v


@pytest.mark.gpu
def test_staged_scans_match_unstaged_at_full_size(gpu):
    """Variant 19 stages the interaction's candidate chunks through LDS from
    524,288 material particles per launch (the 1.55 M-particle column has
    1,024,000), integrates split (the half step beside the step's start, no
    ini copy) and runs Adami and Shepard in other shapes; variant 18 is the
    same scan unstaged with the reference's ini copy.  On a perturbed column,
    over 22 steps (kept tables, a re-sort and a Shepard filter at step 20),
    every result bit agrees."""
    from loquat_amd import abi
    c = cases.perturb(cases.column_3d(nsteps=22), seed=61)
    assert c.np - c.derived().npb >= 1 << 19 and c.par.stress_shepard_filter == 20
    res = []
    for v in (18, 19):
        abi.check(abi.lib().lq_set_interaction_variant(v))
        e, (rid,) = _engine([c], 22)
        e.step(22)
        res.append(e.download(rid))
        e.close()
    abi.check(abi.lib().lq_set_interaction_variant(abi.DEFAULT_INTERACTION_VARIANT))
    assert _same(res[0], res[1])


@pytest.mark.gpu
def test_autotune_picks_a_variant_and_keeps_results(gpu):
    c = cases.perturb(cases.small_column_3d(seed=51), seed=52)
    e, (rid,) = _engine([c], 6)
    best, ms = e.autotune(2)
    # the fastest, or the engine's default variant when that is within 1 %
    assert best in range(abi.INTERACTION_VARIANTS) and all(m > 0 for m in ms)
    assert ms[best] == min(ms) or (best == abi.DEFAULT_INTERACTION_VARIANT and ms[best] <= 1.01 * min(ms))
    e.step(6)
    tuned = e.download(rid)
    e.close()
    e, (rid,) = _engine([c], 6)
    e.step(6)
    plain = e.download(rid)
    e.close()
    for k in ("posd", "velrhop", "str"):
        assert rel_err(tuned[k], plain[k]) < 1e-5, k


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_slab_trajectory.json")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="oracle full-run fixture not generated")
def test_c1_slab_full_run_matches_oracle(gpu):
    """Plane-strain sand column (8-particle slab between side walls), 20,000
    steps on the GPU: runout, centre of mass, height and free-surface profile
    within 1 % of the FP64 oracle's full run (tests/golden/c1_slab_trajectory.json,
    written by tests/golden/make_c1_trajectory.py)."""
    import json

    from golden.metrics import collapse_metrics

    gold = json.load(open(GOLDEN))
    case = cases.column_2d_slab(nsteps=gold["steps"])
    e, (rid,) = _engine([case])
    e.step(gold["steps"])
    got = collapse_metrics(case, e.download(rid)["posd"])
    e.close()
    h0 = case.meta["col"][2]
    assert got["below_floor"] == 0
    assert abs(got["runout"] - gold["runout"]) <= 0.01 * gold["runout"], (got["runout"], gold["runout"])
    assert abs(got["com_x"] - gold["com_x"]) <= 0.01 * gold["com_x"], (got["com_x"], gold["com_x"])
    assert abs(got["height"] - gold["height"]) <= 0.01 * h0, (got["height"], gold["height"])
    # free surface: bins populated in both, within 1 % of the initial column height
    n = min(len(got["profile"]), len(gold["profile"]))
    diffs = [abs(a - b) for a, b in zip(got["profile"][:n], gold["profile"][:n]) if a is not None and b is not None]
    assert len(diffs) >= 0.8 * n
    assert max(diffs) <= 0.01 * h0, max(diffs)


GOLDEN_C2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_column_trajectory.json")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(GOLDEN_C2), reason="oracle full-run fixture not generated")
def test_c2_column_full_run_matches_oracle(gpu):
    """The benchmark's 1.55 M-particle 3D column, all 5,000 steps on the GPU
    (default engine settings: a neighbour search every 12 steps, z-binned
    cells, the default interaction variant): runout, centre of mass, height and free-surface profile within
    1 % of the FP64 oracle's full run (tests/golden/c2_column_trajectory.json,
    written by tests/golden/make_c2_trajectory.py)."""
    import json

    from golden.metrics import collapse_metrics

    gold = json.load(open(GOLDEN_C2))
    case = cases.column_3d(nsteps=gold["steps"])
    e, (rid,) = _engine([case])
    e.step(gold["steps"])
    assert e.info(rid)["steps"] == gold["steps"]
    got = collapse_metrics(case, e.download(rid)["posd"])
    e.close()
    h0 = case.meta["col"][2]
    print("\nC2 full run, GPU vs oracle:", {k: (round(got[k], 5), round(gold[k], 5))
                                          for k in ("runout", "com_x", "com_z", "height")})
    assert got["below_floor"] == 0 and got["count"] == gold["count"]
    assert abs(got["runout"] - gold["runout"]) <= 0.01 * gold["runout"], (got["runout"], gold["runout"])
    assert abs(got["com_x"] - gold["com_x"]) <= 0.01 * gold["com_x"], (got["com_x"], gold["com_x"])
    assert abs(got["com_z"] - gold["com_z"]) <= 0.01 * gold["com_z"], (got["com_z"], gold["com_z"])
    assert abs(got["height"] - gold["height"]) <= 0.01 * h0, (got["height"], gold["height"])
    n = min(len(got["profile"]), len(gold["profile"]))
    diffs = [abs(a - b) for a, b in zip(got["profile"][:n], gold["profile"][:n]) if a is not None and b is not None]
    assert len(diffs) >= 0.8 * n
    assert max(diffs) <= 0.01 * h0, max(diffs)


GOLDEN_C3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_slope_trajectory.json")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(GOLDEN_C3), reason="oracle full-run fixture not generated")
def test_c3_slope_full_run_matches_oracle(gpu):
    """The von Wolffersdorff embankment at reduced resolution (323,843
    particles), 3,000 steps on the GPU: runout, centre of mass, height and
    free-surface profile within 1 % of the FP64 oracle's run, and the
    displacement field's mean and 99th percentile within 1 % of the oracle's
    (tests/golden/c3_slope_trajectory.json, written by
    tests/golden/make_c3_trajectory.py)."""
    import json

    from golden.metrics import collapse_metrics, displacement_metrics

    gold = json.load(open(GOLDEN_C3))
    case = cases.slope_vw(npm_target=gold["npm_target"], nsteps=gold["steps"])
    e, (rid,) = _engine([case])
    e.step(gold["steps"])
    assert e.info(rid)["steps"] == gold["steps"]
    posd = e.download(rid)["posd"]
    e.close()
    got = collapse_metrics(case, posd)
    got.update(displacement_metrics(case, posd))
    keys = ("runout", "com_x", "com_z", "height", "disp_mean", "disp_p99")
    print("\nC3 slope full run, GPU vs oracle:", {k: (round(got[k], 6), round(gold[k], 6)) for k in keys})
    h0 = 1.0   # embankment height, m
    assert got["below_floor"] == 0 and got["count"] == gold["count"]
    for k in ("runout", "com_x", "com_z", "disp_mean", "disp_p99"):
        assert abs(got[k] - gold[k]) <= 0.01 * abs(gold[k]), (k, got[k], gold[k])
    assert abs(got["height"] - gold["height"]) <= 0.01 * h0, (got["height"], gold["height"])
    n = min(len(got["profile"]), len(gold["profile"]))
    diffs = [abs(a - b) for a, b in zip(got["profile"][:n], gold["profile"][:n]) if a is not None and b is not None]
    assert len(diffs) >= 0.8 * n
    assert max(diffs) <= 0.01 * h0, max(diffs)


def _engine_ordered(case, zbits, steps):
    from loquat_amd.device import Engine
    e = Engine(case.np, case.derived().ncell, max_runs=1)
    e.set_cell_order(zbits)
    return e, e.add(case, steps)


@pytest.mark.gpu
@pytest.mark.parametrize("which,zbits", [("small", 0), ("small", 4), ("c2", 4), ("slope", 4)])
def test_engine_neighbour_counts_exact(gpu, which, zbits):
    """Pruned neighbour scans (column skip, and the z-window over z-binned
    cells when zbits > 0) find exactly the oracle's neighbours, particle by
    particle, on the engine's own moved state (c2: the 1.55 M-particle
    benchmark column; slope: the von Wolffersdorff embankment at 0.3 M
    particles, an inclined free surface)."""
    if which == "small":
        case = cases.perturb(cases.small_column_3d(seed=61), seed=62)
    elif which == "slope":
        case = cases.perturb(cases.slope_vw(npm_target=2e5, nsteps=10), seed=64, v_amp=0.3)
    else:
        case = cases.perturb(cases.column_3d(nsteps=10), seed=63, v_amp=0.3)
    e, rid = _engine_ordered(case, zbits, 4)
    e.step(3)
    got = e.neighbour_counts(rid)
    h = e.download(rid)
    e.close()
    p = case.derived()
    st = O.OState.from_arrays(h["posd"], h["pospres"], h["velrhop"], h["idep"], h["str"], h["stn"])
    O.initial_sort(p, st)
    g = O.grid_build(p, st)
    _, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    ref = np.zeros_like(nbr)
    ref[st.orig] = nbr
    npb = p.npb
    assert (got[:npb] == 0).all() and got[npb:].min() > 0
    assert np.array_equal(got[npb:], ref[npb:])


@pytest.mark.gpu
def test_cell_order_changes_only_sum_order(gpu):
    """z-binned and reference in-cell order: same neighbour sets, so the
    trajectories differ only by FP32 summation order."""
    case = cases.perturb(cases.small_column_3d(seed=71, shepard=2), seed=72)
    res = []
    for zb in (0, 4):
        e, rid = _engine_ordered(case, zb, 5)
        e.step(5)
        res.append(e.download(rid))
        e.close()
    x0 = case.posd
    assert rel_err(res[0]["posd"] - x0, res[1]["posd"] - x0) < 1e-5
    for k in ("velrhop", "str"):
        assert rel_err(res[0][k], res[1][k]) < 1e-5, k


@pytest.mark.gpu
def test_restart_from_output_continues_the_run(gpu, tmp_path):
    """Run 8 steps, write an output, read it back as a new project, run 8
    more: the reloaded state equals the written one bit for bit, and the
    continued run matches the uninterrupted 16-step run (the restart re-sorts
    from input order, so only FP32 sum order differs).  No Shepard filter: its
    cadence counts steps from the start of each launch, as the reference's
    step counter does (src/loquat.cu:307,401), so a restart shifts it."""
    case = cases.perturb(cases.small_column_3d(seed=81), seed=82)
    whole, (rid,) = _engine([case], 16)
    whole.step(16)
    ref = whole.download(rid)
    whole.close()

    e, (rid,) = _engine([case], 16)
    e.step(8)
    mid = e.download(rid)
    e.close()
    par = case.par.copy()
    par.output_number = 1
    abi.save_output(str(tmp_path), par, case.vw, mid, mid["posd"])
    p2, vw2 = abi.load_parameters(str(tmp_path / "output_1"))
    arrs = abi.load_particles(str(tmp_path / "1"), p2)
    posd = abi.load_restart(str(tmp_path / "1"), p2, arrs["idep"])
    for k in ("pospres", "velrhop", "str", "stn"):
        assert np.array_equal(arrs[k], mid[k]), k
    assert np.array_equal(posd, mid["posd"])
    resumed = cases.Case("resumed", p2, vw2, arrs["pospres"], arrs["velrhop"], arrs["idep"], arrs["str"],
                         arrs["stn"], posd, 8)
    e, (rid,) = _engine([resumed], 8)
    e.step(8)
    got = e.download(rid)
    e.close()
    x0 = case.posd
    assert rel_err(got["posd"] - x0, ref["posd"] - x0) < 1e-5
    for k in ("velrhop", "str"):
        assert rel_err(got[k], ref[k]) < 1e-5, k


@pytest.mark.gpu
def test_dynamic_time_step_matches_oracle(gpu):
    """dynamic-time-step = 1: the engine's per-step dt (device max-reductions
    of |v| and |a|) and its trajectory follow the oracle's rule."""
    case = gentle_case(1, seed=91)
    case.par.dynamic_time_step = 1
    case.par.stress_shepard_filter = 5
    nsteps = 20
    e, (rid,) = _engine([case], nsteps)
    e.step(nsteps)
    got = e.download(rid)
    info = e.info(rid)
    e.close()
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    dts = O.run_dt(p, case.vw, st, nsteps)
    ref = O.unpermute(st)
    assert len(set(dts.tolist())) > 1                        # the step did change
    assert info["steps"] == nsteps
    assert info["time"] - p.time == pytest.approx(dts.sum(), rel=1e-6)
    m = slice(p.npb, None)
    x0 = case.posd
    for a, k in enumerate(("x", "y", "z")):
        assert rel_err(got["posd"][m, a] - x0[m, a], getattr(ref, k)[m] - x0[m, a]) < 1e-4, k
    for a, k in enumerate(("vx", "vy", "vz")):
        assert rel_err(got["velrhop"][m, a], getattr(ref, k)[m]) < 1e-4, k


@pytest.mark.gpu
def test_dynamic_run_ends_at_time_max(gpu):
    """A run added with nsteps = 0 and a dynamic step ends on the device once
    time >= time-max (src/loquat.cu:322), batched with a fixed-step run that
    keeps going."""
    case = gentle_case(1, seed=93)
    case.par.dynamic_time_step = 1
    p = case.derived()
    case.par.time_max = p.time + 12.5 * p.dt
    other = gentle_case(1, seed=94)
    e, _ = _engine([], spare=(case, other))
    rid = e.add(case, 0)        # 0: run until time-max
    oid = e.add(other, 40)
    e.step(40)
    info, oinfo = e.info(rid), e.info(oid)
    e.close()
    assert oinfo["steps"] == 40 and oinfo["finished"]
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    dts = O.run_dt(p, case.vw, st, 40)
    t = p.time + np.cumsum(dts)
    expect = int(np.argmax(t >= case.par.time_max)) + 1
    assert info["finished"] and info["steps"] == expect
    assert info["time"] >= case.par.time_max


@pytest.mark.gpu
def test_compact_removes_particles_outside_the_grid(gpu):
    """Material particles outside the grid are removed by lq_engine_compact;
    every other particle ends bit-identical to the same run without
    compaction (no neighbour search ever finds the removed ones)."""
    case = cases.perturb(cases.small_column_3d(seed=101), seed=102)
    p = case.derived()
    gone = p.npb + np.arange(5, 5 + 7 * 37, 37)           # seven material particles
    case.posd[gone, 2] = p.domain_size_z + 0.5              # above the grid
    case.pospres[gone, 2] = case.posd[gone, 2].astype(np.float32)
    e, (rid,) = _engine([case], 12)
    e.step(12)
    ref = e.download(rid)
    e.close()
    e, (rid,) = _engine([case], 12)
    e.step(4)
    assert e.compact() == len(gone)
    assert e.totals()["np"] == case.np - len(gone)
    e.step(8)
    got = e.download(rid)
    info = e.info(rid)
    e.close()
    assert info["steps"] == 12 and info["np"] == case.np
    keep = np.ones(case.np, bool)
    keep[gone] = False
    assert np.isnan(got["pospres"][gone, :3]).all() and np.isnan(got["posd"][gone]).all()
    for k in ("posd", "pospres", "velrhop", "idep", "str", "stn"):
        assert np.array_equal(got[k][keep], ref[k][keep]), k


@pytest.mark.gpu
def test_engine_edge_cases_and_errors(gpu):
    """A walls-only run beside a normal one, a run of zero steps, capacity and
    id errors, and stepping an engine whose runs have all finished."""
    from loquat_amd.device import Engine
    normal = cases.perturb(cases.small_column_3d(seed=111), seed=112)
    walls = cases.small_column_3d(seed=113)
    npb = walls.derived().npb
    keep = slice(0, npb)
    walls = cases.Case("walls-only", walls.par.copy(), walls.vw, walls.pospres[keep].copy(),
                       walls.velrhop[keep].copy(), walls.idep[keep].copy(), walls.str[keep].copy(),
                       walls.stn[keep].copy(), walls.posd[keep].copy(), 5)
    walls.par.np, walls.par.npm = npb, 0
    alone, (ra,) = _engine([normal], 5)
    alone.step(5)
    ref = alone.download(ra)
    alone.close()
    e = Engine(normal.np + walls.np, normal.derived().ncell + walls.derived().ncell, max_runs=2)
    rw = e.add(walls, 5)
    rn = e.add(normal, 5)
    with pytest.raises(abi.LqError) as err:       # run table full
        e.add(normal, 5)
    assert err.value.status == 8
    e.step(9)                                     # runs stop at their own 5 steps
    assert e.info(rn)["steps"] == 5 and e.info(rn)["finished"]
    assert e.info(rw)["steps"] == 5
    assert _same(e.download(rn), ref)
    got_w = e.download(rw)
    assert np.array_equal(got_w["posd"], walls.posd)      # walls never move
    e.step(3)                                     # nothing left to run
    assert e.info(rn)["steps"] == 5
    with pytest.raises(abi.LqError) as err:
        e.info(12345)
    assert err.value.status == 1
    e.remove(rw)
    assert e.totals()["nruns"] == 1
    assert _same(e.download(rn), ref)
    e.close()
    small = Engine(10, 10, 1)
    with pytest.raises(abi.LqError) as err:       # more particles than the engine holds
        small.add(normal, 1)
    assert err.value.status == 8
    small.close()


@pytest.mark.gpu
def test_run_moving_beyond_the_table_reach_is_stopped(gpu):
    """A run whose particles move more than 0.45 cell in half a step (a
    uniform 1000 m/s translation: no strain rate, so the state stays finite)
    is flagged failed by the corrector, because the corrector half's reuse of
    the cell table would no longer be exact; a run beside it is unaffected."""
    import ctypes as C
    good = gentle_case(1, seed=71)
    fast = gentle_case(1, seed=72)
    npb = fast.par.npb
    fast.velrhop[npb:, 0] = 1000.0
    p = fast.derived()
    assert 1000.0 * 0.5 * p.dt > 0.45 * (2.0 / p.i_h)
    e, (rg, rf) = _engine([good, fast], 5)
    e.step(5)
    e.sync()
    lib = abi.lib()
    nfailed = C.c_int32()
    abi.check(lib.lq_engine_poll(e.h, C.byref(nfailed)))
    fg, ff = C.c_int32(), C.c_int32()
    abi.check(lib.lq_engine_run_failed(e.h, rg, C.byref(fg)))
    abi.check(lib.lq_engine_run_failed(e.h, rf, C.byref(ff)))
    assert nfailed.value == 1 and ff.value == 1 and fg.value == 0
    got = e.download(rg)
    e.close()
    e2, (r2,) = _engine([good], 5)
    e2.step(5)
    assert _same(got, e2.download(r2))
    e2.close()


@pytest.mark.gpu
def test_slope_von_wolffersdorff_stays_stable(gpu):
    """The von Wolffersdorff embankment at 0.3 M particles runs 3,000 steps
    (0.105 s; without the model's tension cut and void-ratio bounds this
    embankment blows up at its toe near 0.035 s) without a failure flag:
    finite state, void ratio inside [e_d0, e_i0], no tensile mean stress,
    particles slower than 5 m/s."""
    import ctypes as C
    from loquat_amd import abi
    case = cases.slope_vw(npm_target=2e5, nsteps=3000)
    e, (rid,) = _engine([case], 3000)
    e.step(3000)
    e.sync()
    nfailed = C.c_int32()
    abi.check(abi.lib().lq_engine_poll(e.h, C.byref(nfailed)))
    info = e.info(rid)
    h = e.download(rid)
    e.close()
    assert nfailed.value == 0 and info["finished"] and info["steps"] == 3000
    m = slice(case.par.npb, None)
    vel, s, ev = h["velrhop"][m], h["str"][m], h["idep"][m, 2]
    assert np.isfinite(h["posd"][m]).all() and np.isfinite(vel).all() and np.isfinite(s).all()
    assert np.abs(vel[:, :3]).max() < 5.0
    assert (s[:, :3].sum(1) <= 0).all()
    assert ev.min() >= case.vw.ed0 * (1 - 1e-6) and ev.max() <= case.vw.ei0 * (1 + 1e-6)


@pytest.mark.gpu
def test_autotune_cell_order_keeps_neighbour_sets(gpu):
    """The cell-order tuner returns one of its candidates with a positive time
    for each, leaves the engine in that order, and the neighbour sets after it
    are still exactly the oracle's; the run then steps normally."""
    from loquat_amd import abi
    case = cases.perturb(cases.small_column_3d(seed=81), seed=82)
    e, rid = _engine_ordered(case, 4, 6)
    bits, ms = e.autotune_cell_order(2)
    assert 0 <= bits < abi.CELL_ORDER_CANDIDATES and len(ms) == abi.CELL_ORDER_CANDIDATES
    assert all(m > 0 for m in ms)
    # the fastest, or the starting order (4) when within 1 % of the fastest
    assert bits == int(np.argmin(ms)) or (bits == 4 and ms[4] <= 1.01 * min(ms))
    got = e.neighbour_counts(rid)
    h = e.download(rid)
    p = case.derived()
    st = O.OState.from_arrays(h["posd"], h["pospres"], h["velrhop"], h["idep"], h["str"], h["stn"])
    O.initial_sort(p, st)
    g = O.grid_build(p, st)
    _, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    ref = np.zeros_like(nbr)
    ref[st.orig] = nbr
    assert np.array_equal(got[p.npb:], ref[p.npb:])
    e.step(6)
    e.sync()
    assert e.info(rid)["steps"] == 6 and e.info(rid)["finished"]
    assert np.isfinite(e.download(rid)["velrhop"]).all()
    e.close()


@pytest.mark.gpu
def test_autotune_column_order_keeps_results(gpu):
    """The column-order tuner times its five candidates, pins one of them and
    leaves results bit-identical to an untuned engine's."""
    c = cases.perturb(cases.small_column_3d(seed=61), seed=62)
    e, (rid,) = _engine([c], 6)
    rows, ms = e.autotune_column_order(2)
    assert rows in (0, 2, 4, 8, 16) and e.column_order() == rows and len(ms) == 5 and all(m > 0 for m in ms)
    e.step(6)
    tuned = e.download(rid)
    e.close()
    e, (rid,) = _engine([c], 6)
    e.step(6)
    assert _same(tuned, e.download(rid))
    e.close()
