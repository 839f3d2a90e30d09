"""Pins the FP64 oracle with checks it cannot pass by restating its own
formulas: brute force on small inputs, closed forms that hold exactly on a
cubic lattice, conservation laws, published model properties.

Lattice identity used below: on a cubic lattice the interior sum
M = sum_j V_j (x_j - x_i) (x) grad_i W_ij is isotropic by symmetry, M = c I
(c ~ 1 is the kernel's discrete normalisation), so for any linear field
f = f0 + K x the SPH gradients are exactly c K.
"""
import math

import numpy as np
import pytest

from loquat_amd import abi, cases
from oracle import lq_oracle as O
from helpers import fma32, r2_fp32


def lattice_case(nx=14, ny=14, nz=14, walls=True):
    c = cases.small_column_3d(nx=nx, ny=ny, nz=nz, jitter=0.0, seed=0)
    if not walls:
        npb = c.par.npb
        for k in ("pospres", "velrhop", "idep", "str", "stn", "posd"):
            setattr(c, k, np.ascontiguousarray(getattr(c, k)[npb:]))
        c.par.np -= npb
        c.par.npb = 0
    return c


def sorted_oracle(case):
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    g = O.grid_build(p, st)
    return p, st, g


def interior(p, st, margin_h=2.05):
    """Material particles whose 2h ball lies inside the material block."""
    m = np.arange(st.n) >= p.npb
    xs = [st.x, st.y, st.z]
    mask = m.copy()
    h2 = margin_h * float(p.h)
    for a in range(3):
        lo, hi = xs[a][m].min(), xs[a][m].max()
        mask &= (xs[a] > lo + h2) & (xs[a] < hi - h2)
    return np.nonzero(mask)[0]


# ------------------------------------------------------------------ grid
def test_cell_keys_match_formula():
    case = cases.perturb(cases.small_column_3d(seed=3), seed=4)
    case.posd[case.par.npb + 5, 0] = -1.0
    case.posd[case.par.npb + 6, 2] = np.nan
    case.posd[case.par.npb + 7, 1] = 100.0
    p = case.derived()
    st = O.OState.from_case(case)
    keys = O.cell_keys(p, st)
    f = case.posd.astype(np.float32)
    with np.errstate(invalid="ignore"):
        c = np.floor((np.float32(0.5) * f) * np.float32(p.i_h))
    n = np.array([p.cell_num_x, p.cell_num_y, p.cell_num_z])
    inside = np.all((c >= 0) & (c < n), 1)
    ci = np.where(inside[:, None], c, 0).astype(np.int64)
    layer = (np.arange(case.np) >= p.npb).astype(np.int64)
    want = np.where(inside, (ci[:, 0] * n[1] + ci[:, 1]) * n[2] + ci[:, 2] + layer * p.ncell_one_layer,
                    p.ncell - 1)
    assert np.array_equal(keys, want)
    assert (~inside).sum() == 3


def test_sort_is_stable_and_table_brute_force():
    rng = np.random.default_rng(0)
    keys = rng.integers(0, 50, 3000).astype(np.int32)
    perm, srt = O.stable_sort(keys, 50)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    assert np.array_equal(srt, keys[perm])
    beg, end = O.cell_table(srt, 55)
    for c in range(55):
        idx = np.nonzero(srt == c)[0]
        if len(idx):
            assert beg[c] == idx[0] and end[c] == idx[-1] + 1
        else:
            assert beg[c] == -1 and end[c] == -1


def _round_f32_exact(v) -> np.float32:
    """Nearest FP32 to the rational v, ties to even."""
    from fractions import Fraction
    r = np.float32(float(v))
    cands = [np.nextafter(r, np.float32(-np.inf)), r, np.nextafter(r, np.float32(np.inf))]
    best = min(cands, key=lambda x: (abs(Fraction(float(x)) - v), int(np.float32(x).view(np.uint32)) & 1))
    return np.float32(best)


def test_fma32_is_correctly_rounded():
    """The tests' FP32 fused multiply-add (the neighbour predicate's r2):
    (1 + 2^-12)^2 = 1 + 2^-11 + 2^-24 lies halfway between two FP32 values,
    so the unfused product rounds to even, the fused one keeps the 2^-24, and
    a term far below FP64's resolution decides a tie the FP64 sum cannot
    see; random and near-tie triples against exact rational arithmetic."""
    from fractions import Fraction
    a = np.float32(1 + 2.0 ** -12)
    assert np.float32(a * a) - np.float32(1) == np.float32(2.0 ** -11)
    assert fma32(a, a, np.float32(-1.0)) == np.float32(2.0 ** -11 + 2.0 ** -24)
    assert fma32(a, a, np.float32(2.0 ** -80)) == np.float32(1 + 2.0 ** -11 + 2.0 ** -23)
    assert fma32(a, a, np.float32(-(2.0 ** -80))) == np.float32(1 + 2.0 ** -11)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 3000)).astype(np.float32)
    x[2, :1000] = -(x[0, :1000] * x[1, :1000])                 # cancellation
    x[2, 1000:2000] = x[2, 1000:2000] * np.float32(2.0 ** -40)  # ties seen only exactly
    got = fma32(x[0], x[1], x[2])
    for i in range(x.shape[1]):
        v = Fraction(float(x[0, i])) * Fraction(float(x[1, i])) + Fraction(float(x[2, i]))
        assert got[i] == _round_f32_exact(v), i


def test_neighbour_counts_equal_brute_force():
    case = cases.perturb(cases.small_column_3d(nx=12, ny=6, nz=14, seed=8), seed=9)
    p, st, g = sorted_oracle(case)
    _, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    f = np.stack([st.x, st.y, st.z], 1).astype(np.float32)
    ih = np.float32(p.i_h)
    for i in range(p.npb, st.n):
        d = f[i] - f
        r2 = r2_fp32(d)
        hit = (np.sqrt(r2) * ih) < np.float32(2.0)
        hit[i] = False
        assert nbr[i] == hit.sum(), i


# ----------------------------------------------------------- interaction
def test_uniform_velocity_gives_zero_rates():
    case = lattice_case()
    case.velrhop[:, :3] = np.array([0.3, -0.2, 0.1], np.float32)
    p, st, g = sorted_oracle(case)
    rt, _ = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    m = slice(p.npb, None)
    for a in (rt.drho, rt.wx, rt.wy, rt.wz, *rt.d):
        assert np.abs(a[m]).max() < 1e-9


@pytest.mark.parametrize("A", [
    np.eye(3),
    np.array([[0.1, -0.7, 0.3], [0.4, -0.2, 0.9], [-0.5, 0.6, 0.25]]),
])
def test_linear_velocity_field_exact_on_lattice(A):
    case = lattice_case(walls=False)
    p, st, g = sorted_oracle(case)
    x = np.stack([st.x, st.y, st.z], 1)
    v = x @ A.T
    st.vx[:], st.vy[:], st.vz[:] = v[:, 0], v[:, 1], v[:, 2]
    rt, _ = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    # normalisation constant from pure expansion
    st2 = st.copy()
    st2.vx[:], st2.vy[:], st2.vz[:] = x[:, 0], x[:, 1], x[:, 2]
    r2, _ = O.interaction(p, st2, g["cell_beg"], g["cell_end"])
    I = interior(p, st)
    assert len(I) > 50
    c = r2.d[0][I]
    assert np.allclose(c, c[0], rtol=1e-10) and abs(c[0] - 1.0) < 0.03
    c = c[0]
    S, W = 0.5 * (A + A.T), 0.5 * (A - A.T)
    want_d = [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]]
    for k in range(6):
        assert np.allclose(rt.d[k][I], c * want_d[k], atol=1e-10 * c), k
    # spin = (W_yx, W_zx, W_zy)
    assert np.allclose(rt.wx[I], c * W[1, 0], atol=1e-10)
    assert np.allclose(rt.wy[I], c * W[2, 0], atol=1e-10)
    assert np.allclose(rt.wz[I], c * W[2, 1], atol=1e-10)
    # continuity: drho = -rho c tr(A)
    assert np.allclose(rt.drho[I], -st.rho[I] * c * np.trace(A), rtol=1e-9, atol=1e-9)


def test_linear_stress_field_momentum_exact_on_lattice():
    """a - g = (c / rho) K . e_z for sigma = sigma0 + z K (v = 0, no viscosity);
    geostatic K_zz = -rho g_z is then in equilibrium up to the normalisation."""
    case = lattice_case(walls=False)
    p, st, g = sorted_oracle(case)
    K = np.array([2.0e3, -1.0e3, -26000.0, 500.0, 700.0, -300.0])  # xx yy zz xy xz yz per metre
    s0 = np.array([-5e3, -6e3, -7e3, 100.0, 200.0, 300.0])
    for k in range(6):
        st.s[k][:] = s0[k] + st.z * K[k]
    st.vx[:] = st.vy[:] = st.vz[:] = 0.0
    rt, _ = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    st2 = st.copy()
    st2.vx[:], st2.vy[:], st2.vz[:] = st.x, st.y, st.z
    r2, _ = O.interaction(p, st2, g["cell_beg"], g["cell_end"])
    I = interior(p, st)
    c = r2.d[0][I][0]
    rho = st.rho[I]
    assert np.allclose(rt.ax[I] - p.acc_x, c * K[4] / rho, rtol=1e-9)   # K_xz
    assert np.allclose(rt.ay[I] - p.acc_y, c * K[5] / rho, rtol=1e-9)   # K_yz
    assert np.allclose(rt.az[I] - p.acc_z, c * K[2] / rho, rtol=1e-9)   # K_zz


def test_momentum_conservation_without_walls():
    case = cases.perturb(lattice_case(walls=False), seed=3, v_amp=0.5)
    p, st, g = sorted_oracle(case)
    rt, _ = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    m = float(p.m)
    f = np.stack([rt.ax - p.acc_x, rt.ay - p.acc_y, rt.az - p.acc_z], 1) * m
    assert np.abs(f.sum(0)).max() < 1e-10 * np.abs(f).sum(0).max()


def _two_particles(u: float):
    dx = 0.005
    posd = np.array([[0.05, 0.05, 0.05], [0.05 + 1.0 * 1.2 * dx, 0.05, 0.05]])
    mat = posd
    c = cases.build_case("pair", mat, np.zeros((0, 3)), dx=dx, domain=(0.1, 0.1, 0.1), rho0=2000.0, model=1,
                         mpar=(1e6, 1e6, 0.5, 0.0, 0.0), cs=30.0, alpha=0.5, g=0.0)
    c.velrhop[0, 0], c.velrhop[1, 0] = u, -u
    return c


def test_borderline_pairs_follow_the_fp32_predicate():
    """Two particles 2h (1 + 5e-8) apart: the FP32 test of the reference
    accepts them (their FP32 coordinates differ by less than 2h), FP64 on the
    FP64 positions rejects them.  The neighbour count follows FP32 and the
    borderline diagnostic reports the pair from both sides; 1e-6 closer, both
    agree."""
    for rel, flips in ((5e-8, 2), (-1e-6, 0)):
        c = _two_particles(0.0)
        p = c.derived()
        d = (2.0 / float(np.float32(p.i_h))) * (1.0 + rel)
        c.posd[1, 0] = c.posd[0, 0] + d
        c.pospres[1, 0] = np.float32(c.posd[1, 0])
        st = O.OState.from_case(c)
        g = O.grid_build(p, st)
        O.permute(g["part_idx"], st)
        g = O.grid_build(p, st)
        _, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
        assert list(nbr) == [1, 1]
        assert O.borderline_pairs(p, st, g["cell_beg"], g["cell_end"]) == (flips, 0)


def test_viscosity_dissipates_approach_only():
    for u, approaching in ((0.2, True), (-0.2, False)):
        c = _two_particles(u)
        p = c.derived()
        st = O.OState.from_case(c)
        g = O.grid_build(p, st)
        O.permute(g["part_idx"], st)
        g = O.grid_build(p, st)
        rt, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
        assert list(nbr) == [1, 1]
        i0 = int(np.nonzero(st.orig == 0)[0][0])
        i1 = 1 - i0
        assert rt.ax[i0] == pytest.approx(-rt.ax[i1], rel=1e-12)
        if approaching:
            assert rt.ax[i0] < 0.0 < rt.ax[i1]   # pushes the pair apart
        else:
            assert rt.ax[i0] == 0.0 and rt.ax[i1] == 0.0


# -------------------------------------------------------------- Adami
def _with_artificial_pressure(case, coef: float):
    case.par.artificial_pressure = 1
    case.par.artificial_pressure_coefficient = coef
    return case


def test_artificial_pressure_off_without_tension():
    """R = eps max(-p, 0) / rho^2: with no particle in tension (p >= 0) the
    term adds exactly zero, so the rates equal those with it switched off."""
    case = cases.perturb(lattice_case(walls=False), seed=5, v_amp=0.3)
    rng = np.random.default_rng(6)
    case.pospres[:, 3] = rng.uniform(0.0, 5e3, case.np).astype(np.float32)
    p0, st0, g0 = sorted_oracle(case)
    r0, _ = O.interaction(p0, st0, g0["cell_beg"], g0["cell_end"])
    p1, st1, g1 = sorted_oracle(_with_artificial_pressure(case, 0.3))
    assert p1.kernel_dr > 0.0
    r1, _ = O.interaction(p1, st1, g1["cell_beg"], g1["cell_end"])
    for k in ("ax", "ay", "az", "drho", "wx"):
        assert np.array_equal(getattr(r0, k), getattr(r1, k)), k


def test_artificial_pressure_pair_closed_form():
    """Two particles at rest, zero deviatoric stress, particle 0 in tension:
    the only force is Monaghan's term, a_0 = -m (R_0 + R_1) f^4 grad_0 W with
    R = eps |p| / rho^2 and f = W(r)/W(dr), W the Wendland C2 kernel
    21/(16 pi h^3) (1 - q/2)^4 (2q + 1).  It pushes the pair apart."""
    c = _with_artificial_pressure(_two_particles(0.0), 0.25)
    c.str[:] = 0.0
    c.pospres[:, 3] = (-4.0e4, 0.0)
    p = c.derived()
    st = O.OState.from_case(c)
    g = O.grid_build(p, st)
    O.permute(g["part_idx"], st)
    g = O.grid_build(p, st)
    rt, nbr = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    assert list(nbr) == [1, 1]
    i0 = int(np.nonzero(st.orig == 0)[0][0])
    i1 = 1 - i0
    h = float(p.h)
    r = abs(st.x[i1] - st.x[i0])
    q, qd = r / h, float(p.dr) / h
    shape = lambda q_: (1.0 - 0.5 * q_) ** 4 * (2.0 * q_ + 1.0)
    f = shape(q) / shape(qd)
    alpha = 21.0 / (16.0 * math.pi * h ** 3)
    dwdr = -5.0 * alpha * q * (1.0 - 0.5 * q) ** 3 / h          # dW/dr
    rho = float(st.rho[i0])
    R0 = 0.25 * 4.0e4 / rho ** 2
    grad0 = dwdr * np.sign(st.x[i0] - st.x[i1])                 # x component of grad_0 W
    expect = -float(p.m) * R0 * f ** 4 * grad0
    assert rt.ax[i0] - p.acc_x == pytest.approx(expect, rel=2e-6)
    assert rt.ax[i1] - p.acc_x == pytest.approx(-expect, rel=2e-6)
    assert np.sign(rt.ax[i0]) == np.sign(st.x[i0] - st.x[i1])   # repulsive
    assert rt.ay[i0] == pytest.approx(p.acc_y) and rt.drho[i0] == 0.0


def test_artificial_pressure_conserves_momentum():
    case = _with_artificial_pressure(cases.perturb(lattice_case(walls=False), seed=3, v_amp=0.5), 0.3)
    rng = np.random.default_rng(4)
    case.pospres[:, 3] = rng.uniform(-2e4, 2e4, case.np).astype(np.float32)
    p, st, g = sorted_oracle(case)
    rt, _ = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    m = float(p.m)
    f = np.stack([rt.ax - p.acc_x, rt.ay - p.acc_y, rt.az - p.acc_z], 1) * m
    assert np.abs(f.sum(0)).max() < 1e-10 * np.abs(f).sum(0).max()
    off = case.par.copy()
    off.artificial_pressure = 0
    p0 = abi.derive(off)
    r0, _ = O.interaction(p0, st, g["cell_beg"], g["cell_end"])
    assert np.abs(rt.ax - r0.ax).max() > 1e-2 * np.abs(r0.ax - p.acc_x).max()   # the term is active


def test_adami_extrapolates_linear_hydrostatic_field_exactly():
    case = cases.small_column_3d(seed=12, jitter=0.1)
    p, st, g = sorted_oracle(case)
    rho0, gz = float(p.rho0), float(p.acc_z)
    s_xx, c0 = -4000.0, -1234.0
    st.s[0][:] = s_xx
    st.s[1][:] = 2 * s_xx
    st.s[2][:] = -rho0 * gz * st.z + c0
    st.s[3][:] = 100.0
    st.s[4][:] = -50.0
    st.s[5][:] = 25.0
    walls_before = st.s[:, : p.npb].copy()
    O.adami(p, st, g["cell_beg"], g["cell_end"])
    w = np.arange(p.npb)
    touched = np.any(st.s[:, w] != walls_before, 0)
    assert touched.sum() > 0.1 * p.npb
    t = w[touched]
    assert np.allclose(st.s[2][t], -rho0 * gz * st.z[t] + c0, rtol=1e-10, atol=1e-6)
    assert np.allclose(st.s[0][t], s_xx, rtol=1e-12)
    assert np.allclose(st.s[1][t], 2 * s_xx, rtol=1e-12)
    assert np.allclose(st.s[3][t], 100.0, rtol=1e-12)
    # walls without material neighbours keep their stress
    assert np.array_equal(st.s[:, w[~touched]], walls_before[:, ~touched])


def test_shepard_preserves_uniform_stress():
    case = cases.small_column_3d(seed=14, jitter=0.1)
    p, st, g = sorted_oracle(case)
    for k, v in enumerate((-1e3, -2e3, -3e3, 10.0, 20.0, 30.0)):
        st.s[k][:] = v
    O.shepard(p, st, g["cell_beg"], g["cell_end"])
    for k, v in enumerate((-1e3, -2e3, -3e3, 10.0, 20.0, 30.0)):
        assert np.allclose(st.s[k][p.npb:], v, rtol=1e-12)


# ------------------------------------------------------ constitutive
def dp_params(phi=30.0, c=0.0, psi=0.0, E=5e6, nu=0.3):
    d = cases.dp_constants(phi, c, psi, E, nu)
    par = abi.LqParams()
    par.constitutive_model = 1
    par.model_par_1, par.model_par_2, par.model_par_3, par.model_par_4, par.model_par_5 = (
        d["K"], d["G"], d["kphi"], d["kpsi"], d["kc"])
    return par


def dp_f(s, par):
    sm = (s[0] + s[1] + s[2]) / 3
    dev = s[:3] - sm
    tau = math.sqrt(0.5 * (dev @ dev + 2 * (s[3:] @ s[3:])))
    return tau + float(par.model_par_3) * sm - float(par.model_par_5), sm, tau


def test_dp_elastic_uniaxial_compression():
    par = dp_params()
    K, G = float(par.model_par_1), float(par.model_par_2)
    s0 = np.array([-1e5, -1e5, -1e5, 0, 0, 0], float)
    eps = 1e-6
    s, p, ep, _ = O.constitutive(par, None, 1.0, [0, 0, 0], [0, 0, -eps, 0, 0, 0], s0)
    assert s[2] - s0[2] == pytest.approx(-(K + 4 * G / 3) * eps, rel=1e-9)
    assert s[0] - s0[0] == pytest.approx(-(K - 2 * G / 3) * eps, rel=1e-9)
    assert ep == 0.0 and p == pytest.approx(-np.mean(s[:3]), rel=1e-12)


def test_dp_shear_return_lands_on_cone():
    par = dp_params(phi=30.0, c=2e3, psi=0.0)
    s0 = np.array([-1e5, -1e5, -1e5, 0, 0, 0], float)
    D = np.array([4e-2, -2e-2, -2e-2, 2e-2, 0, 0])
    trial = s0 + 2 * float(par.model_par_2) * D
    assert dp_f(trial, par)[0] > 0.5e5   # the trial is well outside the cone
    s, p, ep, _ = O.constitutive(par, None, 1.0, [0, 0, 0], D, s0)
    f, sm, tau = dp_f(s, par)
    assert abs(f) < 1e-9 * abs(sm)
    # no dilatancy (kpsi = 0): mean stress equals the elastic trial's
    K = float(par.model_par_1)
    assert sm == pytest.approx(-1e5 + K * D[:3].sum(), rel=1e-12)
    assert ep > 0


def test_dp_dilatancy_moves_mean_stress():
    par = dp_params(phi=30.0, c=0.0, psi=10.0)
    s0 = np.array([-1e5, -1e5, -1e5, 0, 0, 0], float)
    D = np.array([6e-2, -3e-2, -3e-2, 0, 0, 0])
    s, _, _, _ = O.constitutive(par, None, 1.0, [0, 0, 0], D, s0)
    f, sm, _ = dp_f(s, par)
    assert abs(f) < 1e-9 * abs(sm)
    # plastic dilation under a prescribed (isochoric) strain increment is
    # balanced by elastic compression: the mean stress becomes more compressive
    assert sm < -1e5


def test_dp_apex_and_tension_cutoff():
    par = dp_params(phi=30.0, c=5e3)
    kc, kphi = float(par.model_par_5), float(par.model_par_3)
    s0 = np.array([1e4, 2e4, 0.0, 3e3, 0, 0], float)   # tensile, beyond the apex
    s, p, ep, _ = O.constitutive(par, None, 1.0, [0, 0, 0], [0] * 6, s0)
    assert np.allclose(s[:3], kc / kphi, rtol=1e-12) and np.allclose(s[3:], 0.0, atol=1e-9)
    assert p == pytest.approx(-kc / kphi) and ep > 0


def _rot(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def _t6(M):
    return np.array([M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]])


def _m33(t):
    return np.array([[t[0], t[3], t[4]], [t[3], t[1], t[5]], [t[4], t[5], t[2]]])


@pytest.mark.parametrize("model", [1, 2, 3])
def test_constitutive_isotropy(model):
    """An isotropic model commutes with rotations: update(R s R^T, R D R^T) =
    R update(s, D) R^T.  A transposed operand or a wrong index breaks this."""
    if model == 1:
        par, vw = dp_params(phi=25.0, c=1e3, psi=5.0), None
    elif model == 2:
        par, vw = abi.LqParams(), None
        par.constitutive_model = 2
        par.model_par_1, par.model_par_2, par.model_par_3, par.model_par_4 = -101.2, -962.1, -877.3, 1229.0
    else:
        par, vw = abi.LqParams(), abi.LqVw(**cases.VW_SAND)
        par.constitutive_model = 3
    S = _m33([-8e4, -5e4, -3e4, 6e3, -4e3, 2e3])
    D = _m33([1e-3, -4e-4, -2e-4, 3e-4, 5e-4, -6e-4])
    R = _rot(model)
    dt = 0.5
    s1, _, _, e1 = O.constitutive(par, vw, dt, [0, 0, 0], _t6(D), _t6(S), voidr=0.7)
    s2, _, _, e2 = O.constitutive(par, vw, dt, [0, 0, 0], _t6(R @ D @ R.T), _t6(R @ S @ R.T), voidr=0.7)
    assert np.allclose(_m33(s2), R @ _m33(s1) @ R.T, rtol=1e-9, atol=1e-9 * np.abs(s1).max())
    assert e1 == pytest.approx(e2, rel=1e-12)


def test_jaumann_rigid_rotation_preserves_invariants():
    """With D = 0 the update is the Jaumann rotation only: to first order in
    dt it is s + dt (W s - s W), which preserves the trace and |s|."""
    par = dp_params(phi=40.0, c=1e7)
    S = np.array([-8e4, -5e4, -3e4, 6e3, -4e3, 2e3])
    w = np.array([0.3, -0.2, 0.5])
    dt = 1e-5
    s, _, _, _ = O.constitutive(par, None, dt, w, [0] * 6, S)
    W = np.array([[0, -w[0], -w[1]], [w[0], 0, -w[2]], [w[1], w[2], 0]])   # spin (W_yx, W_zx, W_zy)
    want = _m33(S) + dt * (W @ _m33(S) - _m33(S) @ W)
    assert np.allclose(_m33(s), want, rtol=1e-10, atol=1e-6)


def test_vw_isotropic_compression_follows_bauer_law():
    """Von Wolffersdorff: a state on the isotropic compression line e = e_i(p)
    stays on it under isotropic compression (e_i = e_i0 exp(-(3p/hs)^n))."""
    vw = abi.LqVw(**cases.VW_SAND)
    par = abi.LqParams()
    par.constitutive_model = 3
    ei = lambda p: vw.ei0 * math.exp(-(3 * p / vw.hs) ** vw.n)
    p0 = 1e4
    s = np.array([-p0, -p0, -p0, 0, 0, 0], float)
    e = ei(p0)
    eps = 1e-5
    for _ in range(20000):
        s, _, _, e = O.constitutive(par, vw, 1.0, [0, 0, 0], [-eps, -eps, -eps, 0, 0, 0], s, voidr=e)
        if -s[0] > 1e6:
            break
    p = -s[0]
    assert p > 1e5
    assert e == pytest.approx(ei(p), rel=2e-3)


def _vw_par():
    par = abi.LqParams()
    par.constitutive_model = 3
    return par, abi.LqVw(**cases.VW_SAND)


def test_vw_tension_cut_gives_zero_stress_like_dp():
    """Von Wolffersdorff under an extension that makes the mean stress
    tensile returns to zero stress, as Drucker-Prager with c = 0 does (its
    tension cut plus apex return); a compressive result is left alone."""
    par, vw = _vw_par()
    S = [-50.0, -40.0, -30.0, 5.0, -3.0, 2.0]
    D = [2e-3, 1e-3, 1.5e-3, 4e-4, 0.0, -2e-4]
    s, p, _, _ = O.constitutive(par, vw, 1.0, [0.1, 0.0, -0.2], D, S, voidr=0.7)
    assert np.all(s == 0.0) and p == 0.0
    sd, _, _, _ = O.constitutive(dp_params(phi=30.0, c=0.0), None, 1.0, [0.1, 0.0, -0.2], D, S)
    assert np.all(sd == 0.0)
    # mild compression stays compressive and is not cut
    s, p, _, _ = O.constitutive(par, vw, 1.0, [0, 0, 0], [-1e-6] * 3 + [0] * 3, S, voidr=0.7)
    assert s[0] + s[1] + s[2] < -120.0 and p > 40.0


def test_vw_void_ratio_held_between_densest_and_loosest():
    """e after the update lies in [e_d0, e_i0] exp(-(3p/h_s)^n) at the new
    mean pressure p (Bauer's law for the densest and loosest states); NaN
    is passed on, not clamped."""
    par, vw = _vw_par()
    bound = lambda e0, s: e0 * math.exp(-(-(s[0] + s[1] + s[2]) / vw.hs) ** vw.n)
    S = [-2e4, -1.5e4, -1e4, 0, 0, 0]
    # strong compaction from the densest state: e would fall below e_d
    s, _, _, e = O.constitutive(par, vw, 1.0, [0, 0, 0], [-0.05, -0.05, -0.05, 0, 0, 0], S, voidr=0.62)
    assert s[0] + s[1] + s[2] < 0 and e == pytest.approx(bound(vw.ed0, s), rel=1e-12)
    # a loose start above e_i is brought down to e_i at the new pressure
    s, _, _, e = O.constitutive(par, vw, 1e-4, [0, 0, 0], [0.0] * 6, S, voidr=1.5)
    assert e == pytest.approx(bound(vw.ei0, s), rel=1e-12)
    # inside the bounds the void ratio follows de = (1 + e) tr(D) dt
    s, _, _, e = O.constitutive(par, vw, 1.0, [0, 0, 0], [-1e-4, 0, 0, 0, 0, 0], S, voidr=0.7)
    assert e == pytest.approx(0.7 - 1.7e-4, rel=1e-12)
    _, _, _, e = O.constitutive(par, vw, 1.0, [0, 0, 0], [0.0] * 6, S, voidr=float("nan"))
    assert math.isnan(e)


def test_vw_stiffness_bounded_outside_the_no_tension_region():
    """Von Wolffersdorff's F has the denominator 2 + sqrt(2) tan(psi)
    cos(3 theta), which can vanish once tan(psi) > sqrt(2) (a state with a
    tensile principal stress).  At such a state -- pressure 905 Pa, shear
    3.3 kPa, denominator -1.6e-4 uncapped -- a strain increment of 1.3e-4
    moves the stress by less than 100 Pa; and approaching uniaxial
    compression (tan(psi) -> sqrt(2), cos(3 theta) = -1) the update is
    continuous."""
    par, vw = _vw_par()
    S = [-2121.4392, 186.51776, -780.41656, 3270.0193, -1282.4542, 696.94556]
    D = [-18.57, 0.4588, -11.64, -4.426, -7.634, 7.994]
    s, _, _, _ = O.constitutive(par, vw, 6.4e-6, [0, 0, 0], D, S, voidr=0.71)
    assert np.abs(np.asarray(s) - S).max() < 100.0
    Du = [-1e-3, 2e-4, 3e-4, 1e-4, 0, 0]
    out = []
    for lam in (1e-2, 1e-3, 1e-4, 1e-5, 0.0):
        S = [-1e4, -1e4 * lam, -1e4 * lam, 0, 0, 0]
        s, _, _, _ = O.constitutive(par, vw, 1e-3, [0, 0, 0], Du, S, voidr=0.7)
        out.append(np.asarray(s) - S)
    assert np.all(np.isfinite(out))
    jumps = [np.abs(out[i + 1] - out[i]).max() for i in range(len(out) - 1)]
    assert max(jumps[1:]) <= jumps[0] and jumps[-1] < 0.05 * np.abs(out[-1]).max()


def test_midpoint_integration_exact_for_constant_acceleration():
    case = lattice_case(nx=6, ny=6, nz=6, walls=False)
    p = case.derived()
    st = O.OState.from_case(case)
    st.vx[:] = 0.3
    x0, vx0 = st.x.copy(), st.vx.copy()
    rt = O.ORates(st.n)
    rt.ax[:] = 2.0
    ini = O.OIni(st.n)
    O.predictor(p, None, st, ini, rt)
    O.corrector(p, None, st, ini, rt)
    dt = float(p.dt)
    assert np.allclose(st.x, x0 + vx0 * dt + 0.5 * 2.0 * dt * dt, rtol=1e-14)
    assert np.allclose(st.vx, vx0 + 2.0 * dt, rtol=1e-14)


def test_dynamic_time_step_rule():
    """dt = cfl * min(h / (cs + max|v|), sqrt(h / max|a|)) over material
    particles: at rest it is the constant step cfl*h/cs; it shrinks with the
    fastest material particle and not with walls; a large acceleration takes
    over; the |a| term is absent at a run's first step."""
    case = cases.small_column_3d(seed=3)
    p = case.derived()
    st = O.OState.from_case(case)
    rt = O.ORates(st.n)
    npb = p.npb
    dt0 = O.dynamic_dt(p, st, rt, False)
    assert dt0 == pytest.approx(p.cfl * p.h / p.cs, rel=1e-7)
    st.vx[:npb] = 100.0                       # walls do not count
    assert O.dynamic_dt(p, st, rt, False) == dt0
    st.vx[npb + 7] = -3.0
    st.vz[npb + 7] = 4.0                      # |v| = 5
    assert O.dynamic_dt(p, st, rt, False) == pytest.approx(p.cfl * p.h / (p.cs + 5.0), rel=1e-7)
    rt.az[npb + 11] = -9.81
    dt_v = O.dynamic_dt(p, st, rt, True)
    assert dt_v == pytest.approx(p.cfl * p.h / (p.cs + 5.0), rel=1e-7)   # gravity: sqrt(h/g) >> h/cs
    big = 4.0 * p.h / (p.h / (p.cs + 5.0)) ** 2                        # sqrt(h/big) = half the velocity bound
    rt.ax[npb + 20] = big
    assert O.dynamic_dt(p, st, rt, True) == pytest.approx(p.cfl * math.sqrt(p.h / big), rel=1e-6)
    assert O.dynamic_dt(p, st, rt, False) == pytest.approx(p.cfl * p.h / (p.cs + 5.0), rel=1e-7)


def test_dynamic_run_steps_shrink_as_the_column_moves():
    """A collapsing column run with dynamic-time-step = 1: every dt is at most
    the constant step, the first equals it (at rest), later ones fall as the
    material speeds up."""
    case = cases.small_column_3d(nx=8, ny=4, nz=12, seed=4)
    case.par.dynamic_time_step = 1
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    dts = O.run_dt(p, case.vw, st, 30)
    dt_const = np.float32(p.cfl * p.h / p.cs)
    assert dts[0] == pytest.approx(dt_const, rel=1e-7)
    assert (dts <= dt_const * (1 + 1e-7)).all()
    assert dts[-1] < dts[0]
