"""Per-call parity of every HIP stage against the FP64 oracle on identical
inputs, through the C ABI.

Bars: grid (cell ids, sort order, cell table) and neighbour counts bit-exact;
fields within 1e-5 of the oracle, normwise per field:
max_i |gpu - oracle| / max_i |oracle|.
"""
import numpy as np
import pytest

from helpers import r2_fp32, rel_err
from loquat_amd import abi, cases
from oracle import lq_oracle as O

TOL = 1e-5
WU = (-101.2, -962.1, -877.3, 1229.0, 0.0)


def _dev():
    from loquat_amd import device
    return device


def rounded_state(st: O.OState):
    """FP32 roundings of a state: the host arrays the GPU gets, and the
    oracle state holding exactly the same values (FP64 positions kept)."""
    n = st.n
    pospres = np.zeros((n, 4), np.float32)
    pospres[:, 0], pospres[:, 1], pospres[:, 2] = st.x, st.y, st.z
    pospres[:, 3] = st.p
    velrhop = np.stack([st.vx, st.vy, st.vz, st.rho], 1).astype(np.float32)
    idep = np.zeros((n, 4), np.float32)
    idep[:, 0], idep[:, 1], idep[:, 2] = st.id, st.ep, st.voidr
    s = st.s.T.astype(np.float32)
    e = st.e.T.astype(np.float32)
    posd = st.posd.copy()
    host = dict(posd=posd, pospres=pospres, velrhop=velrhop, idep=idep, str=s, stn=e, orig=st.orig.copy())
    ost = O.OState.from_arrays(posd, pospres, velrhop, idep, s, e, st.orig)
    return host, ost


def sorted_inputs(case):
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    host, ost = rounded_state(st)
    ds = _dev().DeviceState.from_host(host["posd"], host["pospres"], host["velrhop"], host["idep"], host["str"],
                                      host["stn"], host["orig"])
    return p, ds, ost


def model_case(model: int, seed: int = 3):
    c = cases.perturb(cases.small_column_3d(nx=14, ny=8, nz=20, seed=seed), seed=seed + 1)
    if model == 2:
        c.par.constitutive_model = 2
        (c.par.model_par_1, c.par.model_par_2, c.par.model_par_3, c.par.model_par_4, c.par.model_par_5) = WU
    elif model in (3, "3-bounds"):
        c.par.constitutive_model = 3
        c.vw = abi.LqVw(**cases.VW_SAND)
        rng = np.random.default_rng(seed)
        c.idep[c.par.npb:, 2] = (0.7 * (1 + rng.uniform(-0.05, 0.05, c.npm))).astype(np.float32)
        if model == "3-bounds":
            # a tenth of the material nearly stress-free (the tension cut
            # where the rates extend it), a tenth with e just outside
            # [e_d, e_i] (about [0.58, 1.05] at these pressures).  Small
            # subsets, so that the fields' norms, against which the FP32
            # error is measured, stay those of the ordinary particles.
            k = c.par.npb + rng.permutation(c.npm)
            tenth = c.npm // 10
            low = k[:tenth]
            c.str[low, :3] = -2.0
            c.str[low, 3:] = 0.5
            c.idep[k[tenth:2 * tenth], 2] = rng.choice([0.55, 1.2], tenth).astype(np.float32)
    return c


def cases_for_grid():
    c1 = cases.column_2d_small(nsteps=1)
    s3 = cases.perturb(cases.small_column_3d(seed=5), seed=6)
    # shuffled order inside each block, some material particles outside the grid
    sh = cases.perturb(cases.small_column_3d(seed=7), seed=8)
    rng = np.random.default_rng(9)
    npb, n = sh.par.npb, sh.par.np
    order = np.concatenate([rng.permutation(npb), npb + rng.permutation(n - npb)])
    for k in ("pospres", "velrhop", "idep", "str", "stn", "posd"):
        setattr(sh, k, np.ascontiguousarray(getattr(sh, k)[order]))
    out = npb + rng.choice(n - npb, 25, replace=False)
    sh.posd[out[:10], 0] = -0.05
    sh.posd[out[10:20], 2] = 10.0
    sh.posd[out[20:], 1] = np.nan
    sh.pospres[:, :3] = sh.posd.astype(np.float32)
    return {"c1": c1, "small3d": s3, "shuffled_outside": sh}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1", "small3d", "shuffled_outside"])
def test_grid_bit_exact(gpu, name):
    dev = _dev()
    case = cases_for_grid()[name]
    p = case.derived()
    st = O.OState.from_case(case)
    ref = O.grid_build(p, st)
    ds = dev.DeviceState.from_case(case)
    g = dev.Grid(case.np, p.ncell)
    g.build(p, ds)
    got = g.arrays(case.np, p.ncell)
    for k in ("part_cel", "part_idx", "cell_beg", "cell_end"):
        assert np.array_equal(got[k], ref[k]), k
    # the permute follows part_idx exactly
    dst = dev.DeviceState(case.np)
    g.permute(case.np, ds, dst)
    h = dst.host()
    perm = ref["part_idx"]
    assert np.array_equal(h["orig"], perm)
    assert np.array_equal(h["posd"], case.posd[perm], equal_nan=True)
    assert np.array_equal(h["str"], case.str[perm])
    g.close()


@pytest.mark.gpu
def test_grid_full_size_c2(gpu):
    """Grid of the 1.55 M-particle 3D column: bit-exact against the oracle."""
    dev = _dev()
    case = cases.column_3d(nsteps=1)
    p = case.derived()
    ref = O.grid_build(p, O.OState.from_case(case))
    ds = dev.DeviceState.from_case(case)
    g = dev.Grid(case.np, p.ncell)
    g.build(p, ds)
    got = g.arrays(case.np, p.ncell)
    for k in ("part_cel", "part_idx", "cell_beg", "cell_end"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.all(np.diff(got["part_cel"]) >= 0)
    g.close()


@pytest.mark.gpu
def test_interaction_full_size_c2(gpu):
    """Interaction on the perturbed 1.55 M-particle column (the benchmark's
    size): neighbour counts bit-exact, every rate (accad, spin, acestn)
    within 1e-5 of the oracle, and accad within 1e-5 of each particle's
    absolute term sum.  Prints the errors."""
    dev = _dev()
    import torch
    case = cases.perturb(cases.column_3d(nsteps=1), seed=61, v_amp=0.3)
    p, ds, ost = sorted_inputs(case)
    g, beg, end, ref = _grid_tables(p, ds, ost)
    rt = dev.DeviceRates(ds.n)
    nbr = torch.full((ds.n,), -1, dtype=torch.int32, device="cuda")
    dev.interaction(p, ds, rt, beg, end, nbr)
    ort, onbr = O.interaction(p, ost, ref["cell_beg"], ref["cell_end"])
    h = rt.host()
    m = slice(p.npb, None)
    assert np.array_equal(nbr.cpu().numpy()[m], onbr[m])
    errs = {}
    for a, k in enumerate(("ax", "ay", "az", "drho")):
        errs[k] = rel_err(h["accad"][m, a], getattr(ort, k)[m])
    for a, k in enumerate(("wx", "wy", "wz")):
        errs[k] = rel_err(h["spin"][m, a], getattr(ort, k)[m])
    for c in range(6):
        errs[f"d{c}"] = rel_err(h["acestn"][m, c], ort.d[c, m])
    ts = term_sum_errors(p, h["accad"], ort, ost, ref["cell_beg"], ref["cell_end"])
    g.close()
    print("\nC2 interaction, max-norm relative errors:", {k: f"{v:.2e}" for k, v in errs.items()})
    print("C2 accad, max per-particle error / sum_j |term_ij|:", {k: f"{v:.2e}" for k, v in ts.items()})
    bad = {k: v for k, v in {**errs, **{"termsum " + k: v for k, v in ts.items()}}.items() if not v < TOL}
    assert not bad, bad


@pytest.mark.gpu
def test_full_size_c3_per_field(gpu):
    """The 4.9 M-particle embankment (4.0 M material, von Wolffersdorff with
    the +-5 % void-ratio field; velocities and stresses perturbed so that
    every term acts): grid bit-exact from the unsorted input, then one call of
    each stage from identical inputs -- Adami wall stress, interaction
    (neighbour counts exact), predictor, corrector -- with every field within
    1e-5 of the FP64 oracle (max-norm relative).  Prints the per-field
    errors."""
    dev = _dev()
    import torch
    case = cases.perturb(cases.slope_vw(nsteps=1), seed=71, v_amp=0.3)
    p = case.derived()
    ref0 = O.grid_build(p, O.OState.from_case(case))
    ds0 = dev.DeviceState.from_case(case)
    g0 = dev.Grid(case.np, p.ncell)
    g0.build(p, ds0)
    got = g0.arrays(case.np, p.ncell)
    for k in ("part_cel", "part_idx", "cell_beg", "cell_end"):
        assert np.array_equal(got[k], ref0[k]), k
    g0.close()
    del ds0, ref0, got
    p, ds, ost = sorted_inputs(case)
    g, beg, end, ref = _grid_tables(p, ds, ost)
    npb, m = p.npb, slice(p.npb, None)
    errs = {}
    dev.adami(p, ds, beg, end)
    O.adami(p, ost, ref["cell_beg"], ref["cell_end"])
    h = ds.host()
    for c in range(6):
        errs[f"adami str{c}"] = rel_err(h["str"][:npb, c], ost.s[c, :npb])
    rt = dev.DeviceRates(ds.n)
    nbr = torch.full((ds.n,), -1, dtype=torch.int32, device="cuda")
    dev.interaction(p, ds, rt, beg, end, nbr)
    ort, onbr = O.interaction(p, ost, ref["cell_beg"], ref["cell_end"])
    assert np.array_equal(nbr.cpu().numpy()[m], onbr[m])
    hr = rt.host()
    for a, k in enumerate(("ax", "ay", "az", "drho")):
        errs["interaction " + k] = rel_err(hr["accad"][m, a], getattr(ort, k)[m])
    for a, k in enumerate(("wx", "wy", "wz")):
        errs["interaction " + k] = rel_err(hr["spin"][m, a], getattr(ort, k)[m])
    for c in range(6):
        errs[f"interaction d{c}"] = rel_err(hr["acestn"][m, c], ort.d[c, m])
    for k, v in term_sum_errors(p, hr["accad"], ort, ost, ref["cell_beg"], ref["cell_end"]).items():
        errs[f"accad {k} / term sum"] = v
    # the integrators from the oracle's rates, so that their inputs are identical
    rt, ort = _rates_from_oracle(ort, ds.n)
    before = ost.copy()
    ini, oini = dev.DeviceIni(ds.n), O.OIni(ds.n)
    dev.predictor(p, case.vw, ds, ini, rt)
    O.predictor(p, case.vw, ost, oini, ort)
    errs.update({"predictor " + k: v for k, v in _state_errors(p, ds.host(), ost, before).items()})
    dev.corrector(p, case.vw, ds, ini, rt)
    O.corrector(p, case.vw, ost, oini, ort)
    errs.update({"corrector " + k: v for k, v in _state_errors(p, ds.host(), ost, before).items()})
    g.close()
    print("\nC3 per-field max-norm relative errors:")
    for k, v in errs.items():
        print(f"  {k:24s} {v:.2e}")
    bad = {k: v for k, v in errs.items() if not v < TOL}
    assert not bad, bad


def term_sum_errors(p, accad, ort, ost, beg, end) -> dict:
    """Per-particle error of accad against the oracle relative to the
    particle's own absolute term sum sum_j |term_ij| (+ |g|): near
    equilibrium the pair terms cancel, so the pointwise relative error says
    nothing, while this one measures the FP32 summation against its scale."""
    oabs = O.interaction_abs(p, ost, beg, end)
    m = slice(p.npb, None)
    out = {}
    for a, k in enumerate(("ax", "ay", "az", "drho")):
        err = np.abs(accad[m, a].astype(np.float64) - getattr(ort, k)[m])
        den = getattr(oabs, k)[m]
        zero = den == 0.0
        assert np.all(err[zero] == 0.0), k
        out[k] = float((err[~zero] / den[~zero]).max()) if (~zero).any() else 0.0
    return out


def _grid_tables(p, ds, ost):
    dev = _dev()
    g = dev.Grid(ds.n, p.ncell)
    g.build(p, ds)
    beg, end = g.table_ptrs()
    ref = O.grid_build(p, ost)
    assert np.array_equal(ref["part_idx"], np.arange(ds.n)), "input must already be sorted"
    return g, beg, end, ref


@pytest.mark.gpu
def test_adami_parity(gpu):
    dev = _dev()
    p, ds, ost = sorted_inputs(cases.perturb(cases.small_column_3d(seed=11), seed=12))
    g, beg, end, ref = _grid_tables(p, ds, ost)
    dev.adami(p, ds, beg, end)
    O.adami(p, ost, ref["cell_beg"], ref["cell_end"])
    h = ds.host()
    npb = p.npb
    for c in range(6):
        assert rel_err(h["str"][:npb, c], ost.s[c, :npb]) < TOL, c
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", list(range(abi.INTERACTION_VARIANTS)))
@pytest.mark.parametrize("which", ["small3d", "c1"])
def test_interaction_parity(gpu, variant, which):
    dev = _dev()
    base = cases.small_column_3d(seed=13) if which == "small3d" else cases.column_2d_small(nsteps=1)
    p, ds, ost = sorted_inputs(cases.perturb(base, seed=14))
    g, beg, end, ref = _grid_tables(p, ds, ost)
    abi.check(abi.lib().lq_set_interaction_variant(variant))
    try:
        import torch
        rt = dev.DeviceRates(ds.n)
        nbr = torch.full((ds.n,), -1, dtype=torch.int32, device="cuda")
        dev.interaction(p, ds, rt, beg, end, nbr)
        ort, onbr = O.interaction(p, ost, ref["cell_beg"], ref["cell_end"])
        h = rt.host()
        nb = nbr.cpu().numpy()
    finally:
        abi.check(abi.lib().lq_set_interaction_variant(abi.DEFAULT_INTERACTION_VARIANT))
    npb = p.npb
    assert np.array_equal(nb[npb:], onbr[npb:])
    assert nb[npb:].min() > 0
    m = slice(npb, None)
    assert rel_err(h["accad"][m, 0], ort.ax[m]) < TOL
    assert rel_err(h["accad"][m, 1], ort.ay[m]) < TOL
    assert rel_err(h["accad"][m, 2], ort.az[m]) < TOL
    assert rel_err(h["accad"][m, 3], ort.drho[m]) < TOL
    assert rel_err(h["spin"][m, 0], ort.wx[m]) < TOL
    assert rel_err(h["spin"][m, 1], ort.wy[m]) < TOL
    assert rel_err(h["spin"][m, 2], ort.wz[m]) < TOL
    for c in range(6):
        assert rel_err(h["acestn"][m, c], ort.d[c, m]) < TOL, c
    g.close()


@pytest.mark.gpu
def test_interaction_parity_artificial_pressure(gpu):
    """artificial-pressure = 1 with a third of the particles in tension."""
    dev = _dev()
    import torch
    base = cases.perturb(cases.small_column_3d(seed=15), seed=16)
    base.par.artificial_pressure = 1
    base.par.artificial_pressure_coefficient = 0.3
    rng = np.random.default_rng(17)
    base.pospres[:, 3] = rng.uniform(-1e4, 2e4, base.np).astype(np.float32)
    p, ds, ost = sorted_inputs(base)
    assert p.kernel_dr > 0.0 and (ost.p[p.npb:] < 0).sum() > 100
    g, beg, end, ref = _grid_tables(p, ds, ost)
    rt = dev.DeviceRates(ds.n)
    nbr = torch.full((ds.n,), -1, dtype=torch.int32, device="cuda")
    dev.interaction(p, ds, rt, beg, end, nbr)
    ort, onbr = O.interaction(p, ost, ref["cell_beg"], ref["cell_end"])
    h = rt.host()
    npb = p.npb
    m = slice(npb, None)
    assert np.array_equal(nbr.cpu().numpy()[npb:], onbr[npb:])
    for a, k in enumerate(("ax", "ay", "az", "drho")):
        assert rel_err(h["accad"][m, a], getattr(ort, k)[m]) < TOL, k
    # the term is active: without it the accelerations differ well beyond TOL
    off = p.copy()
    off.artificial_pressure = 0
    o0, _ = O.interaction(off, ost, ref["cell_beg"], ref["cell_end"])
    assert rel_err(o0.az[m], ort.az[m]) > 100 * TOL
    g.close()


def _rates_from_oracle(ort: O.ORates, n: int):
    dev = _dev()
    import torch
    rt = dev.DeviceRates(n)
    acc = np.stack([ort.ax, ort.ay, ort.az, ort.drho], 1).astype(np.float32)
    spin = np.stack([ort.wx, ort.wy, ort.wz, np.zeros(n)], 1).astype(np.float32)
    rt.accad.copy_(torch.from_numpy(acc))
    rt.spin.copy_(torch.from_numpy(spin))
    rt.acestn.copy_(torch.from_numpy(ort.d.T.astype(np.float32).copy()))
    return rt, O.ORates.from_device(rt.host())


def _state_errors(p, h, ost, before) -> dict:
    """Per-field error of the material state (displacements, not absolute
    positions)."""
    m = slice(p.npb, None)
    out = {}
    for a, k in enumerate(("x", "y", "z")):
        out["d" + k] = rel_err(h["posd"][m, a] - before.posd[m, a], getattr(ost, k)[m] - before.posd[m, a])
    for a, k in enumerate(("vx", "vy", "vz", "rho")):
        out[k] = rel_err(h["velrhop"][m, a], getattr(ost, k)[m])
    for c in range(6):
        out[f"str{c}"] = rel_err(h["str"][m, c], ost.s[c, m])
        out[f"stn{c}"] = rel_err(h["stn"][m, c], ost.e[c, m])
    out["ep"] = rel_err(h["idep"][m, 1], ost.ep[m], floor=1e-12)
    out["p"] = rel_err(h["pospres"][m, 3], ost.p[m])
    out["voidr"] = rel_err(h["idep"][m, 2], ost.voidr[m], floor=1e-12)
    return out


def _check_state(p, h, ost, before, tol=TOL):
    for k, v in _state_errors(p, h, ost, before).items():
        assert v < tol, (k, v)


@pytest.mark.gpu
@pytest.mark.parametrize("model", [1, 2, 3, "3-bounds"])
def test_predictor_corrector_parity(gpu, model):
    """One predictor and one corrector from the same rates; "3-bounds" makes
    von Wolffersdorff's tension cut and void-ratio bounds act."""
    dev = _dev()
    case = model_case(model)
    p, ds, ost = sorted_inputs(case)
    g, beg, end, ref = _grid_tables(p, ds, ost)
    ort, _ = O.interaction(p, ost, ref["cell_beg"], ref["cell_end"])
    rt, ort = _rates_from_oracle(ort, ds.n)
    before = ost.copy()
    ini = dev.DeviceIni(ds.n)
    oini = O.OIni(ds.n)
    dev.predictor(p, case.vw, ds, ini, rt)
    O.predictor(p, case.vw, ost, oini, ort)
    h = ds.host()
    _check_state(p, h, ost, before)
    if model == "3-bounds":
        m = slice(p.npb, None)
        tr = ost.s[0, m] + ost.s[1, m] + ost.s[2, m]
        exn = np.exp(-(np.maximum(-tr, 0) / case.vw.hs) ** case.vw.n)
        ev = ost.voidr[m]
        assert (np.abs(ost.s[:, m]).max(0) == 0).sum() > 20          # tension cut
        at = lambda e0: np.isclose(ev, e0 * exn, rtol=1e-12, atol=0).sum()
        assert at(case.vw.ed0) > 20 and at(case.vw.ei0) > 20              # void-ratio bounds
    # corrector from the predicted state and the same rates
    dev.corrector(p, case.vw, ds, ini, rt)
    O.corrector(p, case.vw, ost, oini, ort)
    _check_state(p, ds.host(), ost, before)
    g.close()


@pytest.mark.gpu
def test_shepard_parity(gpu):
    dev = _dev()
    import torch
    p, ds, ost = sorted_inputs(cases.perturb(cases.small_column_3d(seed=17), seed=18))
    g, beg, end, ref = _grid_tables(p, ds, ost)
    cache = torch.zeros(ds.n, 6, dtype=torch.float32, device="cuda")
    walls_before = ds.host()["str"][: p.npb].copy()
    dev.smoothing(p, ds, cache, beg, end)
    O.shepard(p, ost, ref["cell_beg"], ref["cell_end"])
    h = ds.host()
    m = slice(p.npb, None)
    for c in range(6):
        assert rel_err(h["str"][m, c], ost.s[c, m]) < TOL, c
    assert np.array_equal(h["str"][: p.npb], walls_before)
    g.close()


def _brute_counts(p, pos32: np.ndarray, npb: int) -> np.ndarray:
    """Neighbour counts of a fresh grid, by brute force on the FP32
    positions: j != i whose cell (inside the grid) lies in the 27-cell stencil
    of i's cell and r2 < r2_cut; walls count material neighbours only."""
    from scipy.spatial import cKDTree
    ih, r2c = np.float32(p.i_h), np.float32(abi.r2_cut(p))
    cells = np.floor((np.float32(0.5) * pos32) * ih).astype(np.int64)
    nmax = np.array([p.cell_num_x, p.cell_num_y, p.cell_num_z])
    inside = np.all((cells >= 0) & (cells < nmax), axis=1)
    pairs = cKDTree(pos32.astype(np.float64)).query_pairs(r=1.01 * float(np.sqrt(r2c)), output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    d = pos32[i] - pos32[j]
    r2 = r2_fp32(d)
    near = np.all(np.abs(cells[i] - cells[j]) <= 1, axis=1) & (r2 < r2c)
    mat = np.arange(len(pos32)) >= npb
    either = mat[i] | mat[j]
    cnt = np.zeros(len(pos32), np.int64)
    np.add.at(cnt, i[near & inside[j] & either], 1)
    np.add.at(cnt, j[near & inside[i] & either], 1)
    return cnt


def long_case(seed: int = 41):
    """A small block at the far end of a 30 m long grid (over 2,500 cells along
    x), where the FP32 cell coordinates are large and the scans' rounding
    allowance grows with the grid."""
    dx = 0.005
    lat = lambda nx, ny, nz, x0, y0, z0: np.stack(np.meshgrid(
        x0 + dx * (0.5 + np.arange(nx)), y0 + dx * (0.5 + np.arange(ny)), z0 + dx * (0.5 + np.arange(nz)),
        indexing="ij"), -1).reshape(-1, 3)
    x0 = 29.9
    walls = lat(14, 7, 3, x0 - 2 * dx, -2 * dx + 0.02, 0.0)
    mat = lat(10, 3, 6, x0, 0.02 + dx, 3 * dx)
    c = cases.build_case("long", mat, walls, dx=dx, domain=(30.2, 0.08, 0.1), rho0=2000.0, model=1,
                         mpar=(1e6, 1e6, 0.5, 0.0, 0.0), cs=30.0, g=-9.81)
    return cases.perturb(c, seed=seed)


@pytest.mark.gpu
def test_long_grid_interaction_parity(gpu):
    dev = _dev()
    import torch
    base = long_case()
    assert base.derived().cell_num_x > 2000
    p, ds, ost = sorted_inputs(base)
    g, beg, end, ref = _grid_tables(p, ds, ost)
    rt = dev.DeviceRates(ds.n)
    nbr = torch.full((ds.n,), -1, dtype=torch.int32, device="cuda")
    dev.interaction(p, ds, rt, beg, end, nbr)
    ort, onbr = O.interaction(p, ost, ref["cell_beg"], ref["cell_end"])
    h = rt.host()
    m = slice(p.npb, None)
    assert np.array_equal(nbr.cpu().numpy()[m], onbr[m])
    for a, k in enumerate(("ax", "ay", "az", "drho")):
        assert rel_err(h["accad"][m, a], getattr(ort, k)[m]) < TOL, k
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("which,frac", [("column", 0.003), ("column", 0.1), ("column", 0.44), ("long", 0.1)])
def test_stale_table_neighbour_sets_are_fresh(gpu, which, frac):
    """Interaction and Adami on a table built before the particles moved (by
    at most disp = frac * 2h, some exactly disp across cell faces, some out
    of and back into the grid): neighbour counts equal a fresh grid's (brute
    force), rates and wall stresses agree with the fresh-grid calls."""
    dev = _dev()
    import torch
    c = cases.perturb(cases.small_column_3d(seed=31), seed=32) if which == "column" else long_case()
    p0 = c.derived()
    cw = 2.0 / float(p0.i_h)
    disp = frac * cw
    npb = p0.npb
    top = p0.cell_num_z * cw
    # two rows of material particles at the grid's top face: one inside that
    # leaves the grid, one outside that enters it
    sel = np.argsort(c.posd[npb:, 2])[-6:] + npb
    x0, y0 = c.posd[sel[0], 0], c.posd[sel[0], 1]
    for r, k in enumerate(sel):
        c.posd[k] = (x0 + 0.004 * (r % 3), y0, top + (0.3 if r >= 3 else -0.3) * disp)
    c.pospres[sel, :3] = c.posd[sel].astype(np.float32)
    p, dsA, ost = sorted_inputs(c)
    g, beg, end, ref = _grid_tables(p, dsA, ost)
    hA = dsA.host()
    n = dsA.n
    orig = hA["orig"]
    rng = np.random.default_rng(33)
    move = np.zeros((n, 3))
    m = np.arange(npb, n)
    u = rng.normal(size=(len(m), 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    move[m] = u * (rng.uniform(0.0, 1.0, len(m)) * disp)[:, None]
    # up to 300 particles moved exactly disp towards their nearest x face
    fx = (np.float32(0.5) * hA["posd"][:, 0].astype(np.float32)) * np.float32(p.i_h)
    frx = fx - np.floor(fx)
    pick = rng.choice(m, min(300, len(m) // 2), replace=False)
    move[pick] = 0.0
    move[pick, 0] = np.where(frx[pick] < 0.5, -disp, disp)
    k_sel = np.nonzero(np.isin(orig, sel))[0]
    move[k_sel] = 0.0
    move[k_sel, 2] = np.where(hA["posd"][k_sel, 2] < top, disp, -disp)
    posB = hA["posd"] + move
    assert np.linalg.norm(move, axis=1).max() <= disp * (1 + 1e-12)

    def state_b():
        return dev.DeviceState.from_host(posB, hA["pospres"], hA["velrhop"], hA["idep"], hA["str"], hA["stn"],
                                         orig)

    dsB = state_b()
    then = (dsA.xd, dsA.yd, dsA.zd)
    rt_s = dev.DeviceRates(n)
    nbr_s = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    dev.interaction_stale(p, dsB, rt_s, then, disp, beg, end, nbr_s)
    dsBw = state_b()
    nbr_ws = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    dev.adami_stale(p, dsBw, disp, beg, end, nbr_ws)
    # fresh grid of the moved particles
    g2 = dev.Grid(n, p.ncell)
    g2.build(p, dsB)
    dsBs = dev.DeviceState(n)
    g2.permute(n, dsB, dsBs)
    beg2, end2 = g2.table_ptrs()
    rt_f = dev.DeviceRates(n)
    nbr_f = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    dev.interaction(p, dsBs, rt_f, beg2, end2, nbr_f)
    hBs = dsBs.host()
    dev.adami(p, dsBs, beg2, end2)
    hBsw = dsBs.host()
    # fresh order -> stale order, by original index
    pos_of = np.empty(n, np.int64)
    pos_of[orig] = np.arange(n)
    to_b = pos_of[hBs["orig"]]
    bf = _brute_counts(p, posB.astype(np.float32), npb)
    ns, nws = nbr_s.cpu().numpy(), nbr_ws.cpu().numpy()
    nf = np.empty(n, np.int64)
    nf[to_b] = nbr_f.cpu().numpy()
    mat = slice(npb, None)
    assert np.array_equal(nf[mat], bf[mat]), "fresh grid"
    assert np.array_equal(ns[mat], bf[mat]), np.nonzero(ns[mat] != bf[mat])[0][:10]
    assert np.array_equal(nws[:npb], bf[:npb]), np.nonzero(nws[:npb] != bf[:npb])[0][:10]
    assert bf[k_sel].min() > 0  # the rows at the top face have neighbours
    hs, hf = rt_s.host(), rt_f.host()
    for key, w in (("accad", 4), ("spin", 3), ("acestn", 6)):
        f = np.empty_like(hf[key])
        f[to_b] = hf[key]
        for col in range(w):
            assert rel_err(hs[key][mat, col], f[mat, col]) < TOL, (key, col)
    sw = dsBw.host()["str"]
    fw = np.empty_like(hBsw["str"])
    fw[to_b] = hBsw["str"]
    for col in range(6):
        assert rel_err(sw[:npb, col], fw[:npb, col]) < TOL, col
    g.close()
    g2.close()
