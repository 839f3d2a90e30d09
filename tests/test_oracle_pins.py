"""Closed-form pins of the oracle's constitutive updates, integrator half
steps and wall / Shepard stencils.

Each expected value comes from the published law evaluated in a different
representation from the oracle's six-component code: full 3x3 matrices for
the tensor algebra, hand-derived literals for small cases, and the Wendland
C2 kernel written from its textbook definition
W(q) = 21 / (16 pi h^3) (1 - q/2)^4 (2q + 1).  A dropped factor, a swapped
coefficient, a half step taken as a full one, or a stencil that reads the
wall layer instead of the material layer fails at least one test here.

Reference lines pinned:
  Wu rate and guard           src/predictor_corrector.cu:128-164
  Wu / hypoplastic ep         src/predictor_corrector.cu:211, :300
  Drucker-Prager ep           src/predictor_corrector.cu:76-98
  predictor half step         src/predictor_corrector.cu:186-197
  corrector from *_ini        src/predictor_corrector.cu:276-290
  Adami material stencil      src/adami_boundary.cu:38-57, :98-119
  Shepard material stencil    src/field_smoothing.cu:37-56, :100-117
"""
import math

import numpy as np
import pytest

from loquat_amd import abi, cases
from oracle import lq_oracle as O


def _m33(t):
    return np.array([[t[0], t[3], t[4]], [t[3], t[1], t[5]], [t[4], t[5], t[2]]], float)


def _t6(M):
    return np.array([M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]])


def _spin_matrix(w):
    """Spin (W_yx, W_zx, W_zy) as the skew matrix W; the Jaumann term is
    W S - S W (rigid rotation about z with rate Omega has W_yx = Omega)."""
    return np.array([[0.0, -w[0], -w[1]], [w[0], 0.0, -w[2]], [w[1], w[2], 0.0]])


# ------------------------------------------------------------------ Wu
def wu_par(c1, c2, c3, c4, coh=0.0):
    par = abi.LqParams()
    par.constitutive_model = 2
    par.model_par_1, par.model_par_2, par.model_par_3, par.model_par_4, par.model_par_5 = c1, c2, c3, c4, coh
    return par


def wu_rate_matrix(T, D, W, c):
    """Wu-Bauer-Kolymbas (1996): T' = W T - T W + c1 tr(T) D
    + (c2 tr(D) + c3 tr(T D) / tr(T)) T + c4 (T + T*) |D|,
    T* = T - tr(T)/3 I, |D| = sqrt(tr(D D))."""
    trT = np.trace(T)
    Ts = T - trT / 3.0 * np.eye(3)
    nd = math.sqrt(np.trace(D @ D))
    return (W @ T - T @ W + c[0] * trT * D + (c[1] * np.trace(D) + c[2] * np.trace(T @ D) / trT) * T
            + c[3] * (T + Ts) * nd)


def _wu_step(c, s6, d6, w=(0.0, 0.0, 0.0), dt=1e-4, coh=0.0):
    s, _, _, _ = O.constitutive(wu_par(*c, coh=coh), None, dt, list(w), d6, s6)
    return (np.asarray(s) - np.asarray(s6)) / dt


def test_wu_rate_literal_diagonal_case():
    """T = diag(-100, -200, -300), D = diag(1, 0, 0), c = (1, 10, 100, 1000):
    tr T = -600, tr D = 1, tr(TD) = -100, |D| = 1, T* = diag(100, 0, -100),
    so rate = c1 (-600) D + c2 T + c3 T / 6 + c4 diag(0, -200, -400):
    xx = -600 - 1000 - 1666.67, yy = -2000 - 3333.33 - 200000,
    zz = -3000 - 5000 - 400000.  Any two coefficients swapped changes it."""
    got = _wu_step((1.0, 10.0, 100.0, 1000.0), [-100.0, -200.0, -300.0, 0, 0, 0], [1.0, 0, 0, 0, 0, 0], dt=1e-6)
    want = [-600.0 - 1000.0 - 100.0 * 100.0 / 6.0, -2000.0 - 100.0 * 200.0 / 6.0 - 200000.0,
            -3000.0 - 100.0 * 300.0 / 6.0 - 400000.0, 0.0, 0.0, 0.0]
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("k", range(4))
def test_wu_each_term_matches_the_published_rate(k):
    """One coefficient at a time on a non-isotropic stress and strain rate,
    with spin, against the rate evaluated with 3x3 matrices."""
    c = [0.0, 0.0, 0.0, 0.0]
    c[k] = float(np.float32([-101.2, -962.1, -877.3, 1229.0][k]))   # Parameters holds floats
    S6 = [-8e4, -5e4, -3e4, 6e3, -4e3, 2e3]
    D6 = [1e-3, -4e-4, -2e-4, 3e-4, 5e-4, -6e-4]
    w = (0.3, -0.2, 0.5)
    got = _wu_step(c, S6, D6, w=w)
    want = _t6(wu_rate_matrix(_m33(S6), _m33(D6), _spin_matrix(w), c))
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_wu_small_trace_guard_resets_xy_twice_and_keeps_xz():
    """|tr T| < 4 replaces T by diag(-1, -1, -2) with xy and yz zeroed and xz
    kept (src/predictor_corrector.cu:134-143, xy is reset twice).  With only
    c2 and tr D = 1 the rate is that replaced T; with only c4 it is
    (T + T*) |D|, T* = T + 4/3 I."""
    S6 = [-1.0, -0.5, 0.3, 0.4, 0.7, 0.2]          # tr T = -1.2
    D6 = [0.5, 0.25, 0.25, 0.0, 0.0, 0.0]           # tr D = 1, |D| = sqrt(0.375)
    got = _wu_step((0.0, 1.0, 0.0, 0.0), S6, D6)
    assert np.allclose(got, [-1.0, -1.0, -2.0, 0.0, 0.7, 0.0], rtol=1e-12, atol=1e-12)
    got = _wu_step((0.0, 0.0, 0.0, 1.0), S6, D6)
    nd = math.sqrt(0.375)
    want = np.array([-2.0 + 4.0 / 3.0, -2.0 + 4.0 / 3.0, -4.0 + 4.0 / 3.0, 0.0, 1.4, 0.0]) * nd
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # c3 uses tr(T D) / tr T of the replaced T: (-0.5 - 0.25 - 0.5) / -4
    got = _wu_step((0.0, 0.0, 1.0, 0.0), S6, D6)
    assert np.allclose(got, np.array([-1.0, -1.0, -2.0, 0.0, 0.7, 0.0]) * (1.25 / 4.0), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("coh,want", [(0.0, [15.0, 0.0, -15.0]), (3.0, [22.5, 3.0, -16.5])])
def test_wu_mean_stress_capped_at_cohesion(coh, want):
    """s = -10 I, only c1 = -100, D = diag(1, 0.5, 0), dt = 0.01.  The rate
    uses T = s - coh I: c1 tr(T) D = (-100)(-30 - 3 coh) D.  The update
    takes the mean stress above coh and the cap shifts it back to coh
    (src/predictor_corrector.cu:158-164), leaving the deviator."""
    s, _, _, _ = O.constitutive(wu_par(-100.0, 0.0, 0.0, 0.0, coh=coh), None, 0.01, [0, 0, 0],
                                [1.0, 0.5, 0.0, 0, 0, 0], [-10.0, -10.0, -10.0, 0, 0, 0])
    assert np.allclose(s[:3], want, rtol=1e-12, atol=1e-9) and np.all(s[3:] == 0.0)


@pytest.mark.parametrize("model", [2, 3])
def test_hypoplastic_ep_is_equivalent_total_strain(model):
    """Models 2 and 3 set ep = sqrt(2/3 J2(stn)) of the updated total strain
    (src/predictor_corrector.cu:211, :300): stn = (0.01, -0.004, -0.006,
    0.002, 0, 0.001) has J2 = 1.52e-4 + 2 (4e-6 + 1e-6) = 1.62e-4, so
    ep = sqrt(1.08e-4)."""
    if model == 2:
        par, vw = wu_par(-101.2, -962.1, -877.3, 1229.0), None
    else:
        par, vw = abi.LqParams(), abi.LqVw(**cases.VW_SAND)
        par.constitutive_model = 3
    stn = [0.01, -0.004, -0.006, 0.002, 0.0, 0.001]
    _, _, ep, _ = O.constitutive(par, vw, 1e-6, [0, 0, 0], [1e-4, 0, 0, 0, 0, 0], [-8e4, -5e4, -3e4, 0, 0, 0],
                                 stn6=stn, ep=0.5, voidr=0.7)
    assert ep == pytest.approx(math.sqrt(1.08e-4), rel=1e-12)


# ------------------------------------------------------------------- DP
def dp_par(K, G, kphi, kpsi, kc):
    par = abi.LqParams()
    par.constitutive_model = 1
    par.model_par_1, par.model_par_2, par.model_par_3, par.model_par_4, par.model_par_5 = K, G, kphi, kpsi, kc
    return par


def test_dp_tension_cut_ep_literal():
    """Hydrostatic tension 5000 Pa beyond the apex kc/kphi = 2000 Pa, no
    strain: the tension cut adds (5000 - 2000) / K * sqrt(2)/3 to ep and
    lands on the apex (src/predictor_corrector.cu:70-76); no shear return
    follows (tau = 0)."""
    K = 1.0e6
    s, p, ep, _ = O.constitutive(dp_par(K, 4.0e5, 0.5, 0.0, 1000.0), None, 1.0, [0, 0, 0], [0.0] * 6,
                                 [5000.0, 5000.0, 5000.0, 0, 0, 0], ep=0.25)
    assert ep == pytest.approx(0.25 + 3000.0 / K * math.sqrt(2.0) / 3.0, rel=1e-12)
    assert np.allclose(s, [2000.0, 2000.0, 2000.0, 0, 0, 0], rtol=1e-12, atol=1e-9) and p == pytest.approx(-2000.0)


def test_dp_tension_cut_then_shear_return_ep_literal():
    """Tensile state with a deviator: mean 5000 Pa, deviator with
    xx = -yy = 3000 Pa and xy = 1000 Pa, so tau = sqrt(1/2 s:s) =
    sqrt(3000^2 + 1000^2).  The tension cut moves the mean to kc/kphi
    (ep += 3000/K sqrt(2)/3); at the apex f = tau, the shear return with
    kpsi = 0 takes dlambda = tau / G, scales the deviator to zero and adds
    dlambda sqrt(1/3) (src/predictor_corrector.cu:80-98)."""
    K, G = 1.0e6, 4.0e5
    s0 = [8000.0, 2000.0, 5000.0, 1000.0, 0.0, 0.0]
    s, p, ep, _ = O.constitutive(dp_par(K, G, 0.5, 0.0, 1000.0), None, 1.0, [0, 0, 0], [0.0] * 6, s0)
    tau = math.sqrt(3000.0 ** 2 + 1000.0 ** 2)
    assert ep == pytest.approx(3000.0 / K * math.sqrt(2.0) / 3.0 + tau / G * math.sqrt(1.0 / 3.0), rel=1e-12)
    assert np.allclose(s, [2000.0, 2000.0, 2000.0, 0, 0, 0], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("kpsi", [0.0, 0.2])
def test_dp_shear_return_ep_and_mean_stress_closed_form(kpsi):
    """Shear return from an elastic trial computed with 3x3 matrices (Hooke
    with K, G; no spin): f = tau* + kphi sm* - kc > 0, dlambda = f / (G + K
    kphi kpsi), sm = sm* - K kpsi dlambda, deviator scaled to
    (kc - kphi sm) / tau*, ep += dlambda sqrt(1/3 + 2/9 kpsi^2)
    (src/predictor_corrector.cu:80-98)."""
    K, G, kphi, kc = (float(np.float32(v)) for v in (4.1667e6, 1.9231e6, 0.494872, 2.0e3))
    kpsi = float(np.float32(kpsi))
    S0 = _m33([-1e5, -1.2e5, -0.9e5, 5e3, -2e3, 1e3])
    D = _m33([4e-2, -2e-2, -1.5e-2, 2e-2, 1e-2, -5e-3])
    dt = 1.0
    trial = S0 + dt * (2.0 * G * D + (K - 2.0 * G / 3.0) * np.trace(D) * np.eye(3))
    sm_t = np.trace(trial) / 3.0
    dev = trial - sm_t * np.eye(3)
    tau = math.sqrt(0.5 * np.sum(dev * dev))
    f = tau + kphi * sm_t - kc
    assert f > 1e4
    dl = f / (G + K * kphi * kpsi)
    sm = sm_t - K * kpsi * dl
    want = dev * ((kc - kphi * sm) / tau) + sm * np.eye(3)
    s, p, ep, _ = O.constitutive(dp_par(K, G, kphi, kpsi, kc), None, dt, [0, 0, 0], _t6(D), _t6(S0), ep=0.1)
    assert ep == pytest.approx(0.1 + dl * math.sqrt(1.0 / 3.0 + 2.0 / 9.0 * kpsi * kpsi), rel=1e-10)
    assert np.allclose(s, _t6(want), rtol=1e-10, atol=1e-6)
    assert p == pytest.approx(-sm, rel=1e-12)


# ------------------------------------------------------ predictor / corrector
def _one_particle_state():
    st = O.OState(1)
    st.x[0], st.y[0], st.z[0] = 0.125, 0.25, 0.5
    st.vx[0], st.vy[0], st.vz[0], st.rho[0] = 0.3, -0.2, 0.1, 2000.0
    st.p[0], st.ep[0] = 123.0, 0.01
    s0 = [-1.5e4, -2.0e4, -3.0e4, 1.0e3, -2.0e3, 5.0e2]
    e0 = [1e-4, -2e-4, 3e-4, 4e-5, -5e-5, 6e-5]
    for c in range(6):
        st.s[c][0], st.e[c][0] = s0[c], e0[c]
    rt = O.ORates(1)
    rt.ax[0], rt.ay[0], rt.az[0], rt.drho[0] = 2.0, -1.0, -9.81, 50.0
    rt.wx[0], rt.wy[0], rt.wz[0] = 0.4, -0.3, 0.2
    D = [1e-2, -2e-2, 5e-3, 3e-3, -4e-3, 2e-3]
    for c in range(6):
        rt.d[c][0] = D[c]
    return st, rt, np.array(s0), np.array(e0), np.array(D)


def _elastic_rate(s6, D6, w, K, G):
    S, Dm = _m33(s6), _m33(D6)
    W = _spin_matrix(w)
    return _t6(W @ S - S @ W + 2.0 * G * Dm + (K - 2.0 * G / 3.0) * np.trace(Dm) * np.eye(3))


def _elastic_dp():
    par = dp_par(4.0e6, 2.0e6, 0.3, 0.0, 1.0e9)   # cohesion far above any stress here: elastic
    par.np, par.npb, par.npm = 1, 0, 1
    par.dt = 1e-3
    return par


def test_predictor_half_step_closed_form():
    """Predictor (src/predictor_corrector.cu:186-215): v += dt/2 a, x += dt/2
    v_new, rho += dt/2 drho, stn += dt/2 D, and the stress update over dt/2;
    the start-of-step state is copied to ini."""
    par = _elastic_dp()
    dt = float(par.dt)
    st, rt, s0, e0, D = _one_particle_state()
    x0, v0 = np.array([st.x[0], st.y[0], st.z[0]]), np.array([st.vx[0], st.vy[0], st.vz[0]])
    ini = O.OIni(1)
    O.predictor(par, None, st, ini, rt)
    a = np.array([2.0, -1.0, -9.81])
    v1 = v0 + 0.5 * dt * a
    assert np.allclose([st.vx[0], st.vy[0], st.vz[0]], v1, rtol=1e-15)
    assert np.allclose([st.x[0], st.y[0], st.z[0]], x0 + 0.5 * dt * v1, rtol=1e-15)
    assert st.rho[0] == pytest.approx(2000.0 + 0.5 * dt * 50.0, rel=1e-15)
    assert np.allclose(st.e[:, 0], e0 + 0.5 * dt * D, rtol=1e-13, atol=1e-18)
    want_s = s0 + 0.5 * dt * _elastic_rate(s0, D, (0.4, -0.3, 0.2), 4.0e6, 2.0e6)
    assert np.allclose(st.s[:, 0], want_s, rtol=1e-12, atol=1e-9)
    assert st.ep[0] == 0.01 and st.p[0] == pytest.approx(-np.mean(want_s[:3]), rel=1e-12)
    # ini holds the state at the start of the step
    assert ini.rho[0] == 2000.0 and ini.vx[0] == 0.3 and ini.x[0] == 0.125
    assert np.array_equal(ini.s[:, 0], s0) and np.array_equal(ini.e[:, 0], e0)


def test_corrector_full_step_from_ini_closed_form():
    """Corrector (src/predictor_corrector.cu:276-300): every update starts
    from ini with the full dt: v = v_ini + dt a, x = x_ini + dt (v + v_ini)/2,
    rho = rho_ini + dt drho, stn = stn_ini + dt D, stress from str_ini over dt,
    ep from ep_ini -- the state left by the predictor is not read."""
    par = _elastic_dp()
    dt = float(par.dt)
    st, rt, s0, e0, D = _one_particle_state()
    ini = O.OIni(1)
    O.predictor(par, None, st, ini, rt)
    # scramble what the predictor left: the corrector must not read it
    st.rho[0], st.ep[0] = 1.0, 7.0
    for c in range(6):
        st.s[c][0], st.e[c][0] = 1e9, 1e9
    O.corrector(par, None, st, ini, rt)
    v0 = np.array([0.3, -0.2, 0.1])
    v1 = v0 + dt * np.array([2.0, -1.0, -9.81])
    assert np.allclose([st.vx[0], st.vy[0], st.vz[0]], v1, rtol=1e-15)
    assert np.allclose([st.x[0], st.y[0], st.z[0]], np.array([0.125, 0.25, 0.5]) + 0.5 * dt * (v1 + v0), rtol=1e-15)
    assert st.rho[0] == pytest.approx(2000.0 + dt * 50.0, rel=1e-15)
    assert np.allclose(st.e[:, 0], e0 + dt * D, rtol=1e-13, atol=1e-18)
    want_s = s0 + dt * _elastic_rate(s0, D, (0.4, -0.3, 0.2), 4.0e6, 2.0e6)
    assert np.allclose(st.s[:, 0], want_s, rtol=1e-12, atol=1e-9)
    assert st.ep[0] == 0.01


def test_predictor_wu_ep_uses_the_half_step_strain():
    """Model 2 in the predictor: ep = sqrt(2/3 J2(stn + dt/2 D))
    (src/predictor_corrector.cu:197, :211)."""
    par = wu_par(-101.2, -962.1, -877.3, 1229.0)
    par.np, par.npb, par.npm = 1, 0, 1
    par.dt = 1e-3
    st, rt, s0, e0, D = _one_particle_state()
    ini = O.OIni(1)
    O.predictor(par, None, st, ini, rt)
    e1 = _m33(e0 + 0.5e-3 * D)
    assert st.ep[0] == pytest.approx(math.sqrt(2.0 / 3.0 * np.sum(e1 * e1)), rel=1e-12)
    assert st.p[0] == 123.0   # model 2 leaves the pressure as it was


# ------------------------------------------------------------ Adami / Shepard
def wendland(q, h):
    """Wendland C2 in 3-D, support 2h: 21/(16 pi h^3) (1 - q/2)^4 (2q + 1)."""
    return 21.0 / (16.0 * math.pi * h ** 3) * max(1.0 - 0.5 * q, 0.0) ** 4 * (2.0 * q + 1.0)


def _few_particles(mat, walls, rho, stress, acc=(0.0, 0.0, -9.81)):
    dx = 0.005
    c = cases.build_case("few", np.asarray(mat, float), np.asarray(walls, float).reshape(-1, 3), dx=dx,
                         domain=(0.1, 0.1, 0.1), rho0=2000.0, model=1, mpar=(1e6, 1e6, 0.5, 0.0, 0.0), cs=30.0,
                         g=acc[2])
    c.par.acc_x, c.par.acc_y = acc[0], acc[1]
    c.velrhop[:, 3] = np.asarray(rho, np.float32)
    c.str[:] = np.asarray(stress, np.float32)
    p = c.derived()
    st = O.OState.from_case(c)
    g = O.grid_build(p, st)
    O.permute(g["part_idx"], st)
    g = O.grid_build(p, st)
    return p, st, g


def _at(st, orig):
    return int(np.nonzero(st.orig == orig)[0][0])


def test_shepard_three_particle_weighted_average_ignores_walls():
    """Three material particles at mutual distances 0.9h, 1.5h, 1.75h with
    different densities (V = m / rho) and stresses, and one wall particle
    1.2h below the first carrying 1e8 Pa: each material stress becomes
    sum_j s_j W_ij V_j / sum_j W_ij V_j over material j, itself included
    (src/field_smoothing.cu:37-56, :100-117); the wall is not read."""
    h = 0.006
    base = np.array([0.05, 0.05, 0.05])
    mat = [base, base + [0.9 * h, 0.0, 0.0], base + [0.0, 1.5 * h, 0.0]]
    wall = [base + [0.0, 0.0, -1.2 * h]]
    rho = [1.0e3, 2000.0, 2100.0, 1900.0]          # wall first
    S = np.array([[1e8] * 6, [-1e3, -2e3, -3e3, 100, 200, 300], [-5e3, -4e3, -6e3, -50, 70, 10],
                  [-2e4, -1e4, -1.5e4, 400, -300, 250]])
    p, st, g = _few_particles(mat, wall, rho, S)
    hh = float(p.h)
    assert hh == pytest.approx(h, rel=1e-6)
    O.shepard(p, st, g["cell_beg"], g["cell_end"])
    X = np.array(mat)
    V = float(p.m) / np.array(rho[1:], float)
    for i in range(3):
        w = np.array([wendland(np.linalg.norm(X[i] - X[j]) / hh, hh) * V[j] for j in range(3)])
        want = (w[:, None] * S[1:]).sum(0) / w.sum()
        got = st.s[:, _at(st, 1 + i)]
        assert np.allclose(got, want, rtol=2e-6), i
    assert np.all(st.s[:, _at(st, 0)] == np.float32(1e8))   # the wall is left alone


def test_shepard_material_result_independent_of_wall_stress():
    case = cases.perturb(cases.small_column_3d(seed=14, jitter=0.1), seed=2)
    p = case.derived()
    out = []
    for wall_s in (1e9, -1e9):
        st = O.OState.from_case(case)
        st.s[:, : p.npb] = wall_s
        O.initial_sort(p, st)
        g = O.grid_build(p, st)
        O.shepard(p, st, g["cell_beg"], g["cell_end"])
        out.append(O.unpermute(st).s[:, p.npb:])
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[0]).max() < 1e6


def test_adami_two_material_particles_closed_form():
    """One wall particle w with two material neighbours a, b (different
    stress and density) and a second wall particle next to it holding 1e8
    Pa.  The wall stress is the W V-weighted material average, extrapolated
    to w along gravity: s_w = sum_f W V (s_f - rho0 diag(g_k (x_w - x_f)_k))
    / sum_f W V (src/adami_boundary.cu:38-57, :98-119); the other wall is not
    read.  Gravity has all three components so each diagonal term counts."""
    h = 0.006
    w0 = np.array([0.05, 0.05, 0.05])
    mat = [w0 + [0.3 * h, 0.2 * h, 1.1 * h], w0 + [-0.8 * h, 0.4 * h, 1.4 * h]]
    walls = [w0, w0 + [0.9 * h, 0.0, 0.0]]
    rho = [2000.0, 2000.0, 2150.0, 1850.0]
    S = np.array([[7.0, 7.0, 7.0, 7.0, 7.0, 7.0], [1e8] * 6, [-3e3, -4e3, -9e3, 150, -250, 60],
                  [-1.1e4, -6e3, -2.4e4, -400, 300, -90]])
    acc = (1.5, -2.5, -9.81)
    p, st, g = _few_particles(mat, walls, rho, S, acc=acc)
    O.adami(p, st, g["cell_beg"], g["cell_end"])
    hh, rho0 = float(p.h), float(p.rho0)
    gvec = np.array([float(p.acc_x), float(p.acc_y), float(p.acc_z)])
    assert np.allclose(gvec, acc, rtol=1e-7)
    X = np.array(mat)
    V = float(p.m) / np.array(rho[2:], float)
    for wi, xw in enumerate(walls):
        num, den = np.zeros(6), 0.0
        for f in range(2):
            r = xw - X[f]
            q = np.linalg.norm(r) / hh
            if q >= 2.0:
                continue
            kv = wendland(q, hh) * V[f]
            s = S[2 + f].copy()
            s[:3] -= rho0 * gvec * r
            num += kv * s
            den += kv
        got = st.s[:, _at(st, wi)]
        assert den > 0.0
        assert np.allclose(got, num / den, rtol=2e-6, atol=1e-6), wi


def test_adami_hydrostatic_extrapolation_ignores_wall_stress():
    """A linear hydrostatic material field extrapolates exactly to the walls
    whatever the walls held before: the stencil reads material only."""
    case = cases.small_column_3d(seed=12, jitter=0.1)
    p = case.derived()
    rho0, gz = float(p.rho0), float(p.acc_z)
    res = []
    for wall_s in (3e7, -5e6):
        st = O.OState.from_case(case)
        O.initial_sort(p, st)
        g = O.grid_build(p, st)
        st.s[2][:] = -rho0 * gz * st.z - 1234.0
        st.s[0][:] = st.s[1][:] = -4000.0
        st.s[3][:] = st.s[4][:] = st.s[5][:] = 0.0
        walls = np.arange(st.n) < p.npb
        st.s[:, walls] = wall_s
        O.adami(p, st, g["cell_beg"], g["cell_end"])
        touched = st.s[0, walls] != wall_s
        assert touched.sum() > 0.1 * p.npb
        t = np.nonzero(walls)[0][touched]
        assert np.allclose(st.s[2][t], -rho0 * gz * st.z[t] - 1234.0, rtol=1e-10, atol=1e-6)
        assert np.allclose(st.s[0][t], -4000.0, rtol=1e-12)
        res.append(O.unpermute(st).s[:, : p.npb].copy())
    changed = res[0][0] != 3e7
    assert np.array_equal(res[0][:, changed], res[1][:, changed])


def test_interaction_term_sum_bounds_the_rates():
    """lqo_interaction_abs: per particle, sum_j |term_ij| (+ |g| for the
    accelerations) bounds |rate| (triangle inequality), and for a particle
    with one neighbour equals |rate - g| + |g|."""
    case = cases.perturb(cases.small_column_3d(nx=8, ny=4, nz=10, seed=3), seed=4)
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    g = O.grid_build(p, st)
    rt, _ = O.interaction(p, st, g["cell_beg"], g["cell_end"])
    ab = O.interaction_abs(p, st, g["cell_beg"], g["cell_end"])
    m = slice(p.npb, None)
    for k in ("ax", "ay", "az", "drho"):
        assert np.all(getattr(ab, k)[m] >= np.abs(getattr(rt, k)[m]) * (1 - 1e-12)), k
    assert np.all(ab.az[m] > np.abs(rt.az[m] - p.acc_z))
    two = _few_particles([[0.05, 0.05, 0.05], [0.0536, 0.0512, 0.0488]], np.zeros((0, 3)), [2000.0, 2100.0],
                         [[-1e3, -2e3, -3e3, 10, 20, 30], [-4e3, -1e3, -2e3, -5, 7, 9]])
    p2, st2, g2 = two
    st2.vx[:] = [0.1, -0.2]
    r2, nbr = O.interaction(p2, st2, g2["cell_beg"], g2["cell_end"])
    a2 = O.interaction_abs(p2, st2, g2["cell_beg"], g2["cell_end"])
    assert list(nbr) == [1, 1]
    gz = abs(float(p2.acc_z))
    assert np.allclose(a2.az, np.abs(r2.az - p2.acc_z) + gz, rtol=1e-12)
    assert np.allclose(a2.drho, np.abs(r2.drho), rtol=1e-12)
