"""Synthetic project generator: granular columns and embankments on a
lattice with 3 layers of Adami wall particles, shifted into the positive
octant (the reference computes cells from raw coordinates and ignores
domain_min, src/adami_boundary.cu:80-82).

Writes the reference's project files too: <prefix>.xml (flat typed keys,
src/loquat_io.cu:69-201) and <prefix>.dat (22 columns per particle,
src/loquat_io.cu:226).

Configurations:
  column_2d_small   quasi-2D plane-strain sand column, 50 x 100 x 1 at 2 mm
  column_3d         3D column 80 x 80 x 160 at 5 mm (1,024,000 particles)
  slope_vw          30-degree embankment, von Wolffersdorff, ~4 M particles
  ensemble          256 columns of log-uniform 2k..60k particles
  served_mix        Poisson request stream of columns and embankments, 5k..500k
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

LAYERS = 3


@dataclass
class Case:
    name: str
    par: abi.LqParams                 # raw (underived) parameters, as in the XML
    vw: abi.LqVw
    pospres: np.ndarray               # (n, 4) float32
    velrhop: np.ndarray               # (n, 4) float32
    idep: np.ndarray                  # (n, 4) float32
    str: np.ndarray                   # (n, 6) float32
    stn: np.ndarray                   # (n, 6) float32
    posd: np.ndarray                  # (n, 3) float64
    nsteps: int
    meta: dict = field(default_factory=dict)

    @property
    def np(self) -> int:
        return int(self.par.np)

    @property
    def npm(self) -> int:
        return int(self.par.npm)

    def derived(self) -> abi.LqParams:
        return abi.derive(self.par)

    def run_input(self, nsteps: int | None = None) -> abi.LqRunInput:
        ri = abi.LqRunInput()
        ri.par = self.par.copy()
        ri.vw = self.vw
        ri.host = abi.host_particles(self.pospres, self.velrhop, self.idep, self.str, self.stn)
        ri.posd = self.posd.ctypes.data
        ri.nsteps = int(self.nsteps if nsteps is None else nsteps)
        return ri


def dp_constants(phi_deg: float, c: float, psi_deg: float, E: float, nu: float) -> dict:
    """Drucker-Prager cone matched to Mohr-Coulomb in triaxial extension
    (the convention of examples/sand_collapse_plane.xml:20: phi = 30 degrees
    gives DP-kphi = 0.494872 = 2 sqrt(3) sin(phi) / (3 + sin(phi)))."""
    sp, cp = math.sin(math.radians(phi_deg)), math.cos(math.radians(phi_deg))
    spsi = math.sin(math.radians(psi_deg))
    return dict(
        K=E / (3.0 * (1.0 - 2.0 * nu)),
        G=E / (2.0 * (1.0 + nu)),
        kphi=2.0 * math.sqrt(3.0) * sp / (3.0 + sp),
        kpsi=2.0 * math.sqrt(3.0) * spsi / (3.0 + spsi),
        kc=2.0 * math.sqrt(3.0) * c * cp / (3.0 + sp),
    )


def _lattice(nx: int, ny: int, nz: int, x0: float, y0: float, z0: float, dx: float) -> np.ndarray:
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([x0 + (i.ravel() + 0.5) * dx, y0 + (j.ravel() + 0.5) * dx, z0 + (k.ravel() + 0.5) * dx], 1)


def _base_params(**kw) -> abi.LqParams:
    p = abi.LqParams()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def build_case(name: str, mat: np.ndarray, walls: np.ndarray, *, dx: float, domain: tuple, rho0: float,
               model: int, mpar: tuple, cs: float, hdr: float = 1.2, cfl: float = 0.1, alpha: float = 0.25,
               shepard: int = 0, nsteps: int = 100, g: float = -9.81, z_top: float | np.ndarray | None = None,
               k0: float = 0.5, vw: abi.LqVw | None = None, e0: np.ndarray | None = None,
               time_max: float | None = None) -> Case:
    """Assembles a project: walls first (indices [0, npb)), then material.
    z_top (scalar, or one value per material particle): height of the free
    surface above each particle, for the geostatic initial stress (vertical
    rho0 g (z_top - z), horizontal k0 times that)."""
    npb, npm = len(walls), len(mat)
    n = npb + npm
    posd = np.concatenate([walls, mat]).astype(np.float64)
    pospres = np.zeros((n, 4), np.float32)
    pospres[:, :3] = posd.astype(np.float32)
    velrhop = np.zeros((n, 4), np.float32)
    velrhop[:, 3] = rho0
    idep = np.zeros((n, 4), np.float32)
    idep[npb:, 0] = 1.0                 # object id: walls 0, material 1
    if e0 is not None:
        idep[npb:, 2] = e0
    s = np.zeros((n, 6), np.float32)
    if z_top is not None and npm:
        szz = rho0 * g * (z_top - posd[npb:, 2])   # geostatic, compression negative
        s[npb:, 2] = szz
        s[npb:, 0] = k0 * szz
        s[npb:, 1] = k0 * szz
        pospres[npb:, 3] = -(s[npb:, 0] + s[npb:, 1] + s[npb:, 2]) / 3.0
    stn = np.zeros((n, 6), np.float32)
    dt_est = cfl * dx * hdr / cs
    par = _base_params(
        np=n, npb=npb, npm=npm,
        domain_size_x=domain[0], domain_size_y=domain[1], domain_size_z=domain[2],
        dr=dx, hdr=hdr, rho0=rho0, cs=cs, cfl=cfl, constitutive_model=model,
        model_par_1=mpar[0], model_par_2=mpar[1], model_par_3=mpar[2], model_par_4=mpar[3], model_par_5=mpar[4],
        stress_shepard_filter=shepard, dynamic_time_step=0, integration_method=1, viscosity_type=1,
        artificial_viscosity_alpha=alpha, artificial_pressure=0, acc_z=g,
        time_max=time_max if time_max is not None else dt_est * nsteps * 1.0000001,
        time=0.0, dt=dt_est, output_frequency=1e9, output_number=0, output_format=1,
    )
    return Case(name, par, vw if vw is not None else abi.LqVw(), pospres, velrhop, idep, s, stn, posd, nsteps)


def column(*, name: str, dx: float, col: tuple, box: tuple, quasi2d: bool, jitter: float = 0.0, seed: int = 0,
           rho0: float = 2650.0, E: float = 5e6, nu: float = 0.3, phi: float = 30.0, c: float = 0.0,
           psi: float = 0.0, shepard: int = 0, nsteps: int = 100, alpha: float = 0.25,
           z_margin: float = 1.25) -> Case:
    """Granular column (col = width x, depth y, height z) against the x = 0
    wall, on the floor of a box of horizontal size box = (length x, width y).
    Walls: floor + left wall, plus both y side walls in 3D."""
    o = LAYERS * dx
    nx, nz = round(col[0] / dx), round(col[2] / dx)
    ny = 1 if quasi2d else round(col[1] / dx)
    rng = np.random.default_rng(seed)
    oy = 0.0 if quasi2d else o
    mat = _lattice(nx, ny, nz, o, oy, o, dx)
    if jitter:
        mat = mat + rng.uniform(-jitter, jitter, mat.shape) * dx
    nfx = round((o + box[0]) / dx)
    wall_h = round(col[2] / dx)
    parts = []
    if quasi2d:
        parts.append(_lattice(nfx, 1, LAYERS, 0.0, 0.0, 0.0, dx))           # floor
        parts.append(_lattice(LAYERS, 1, wall_h, 0.0, 0.0, o, dx))          # left wall
        dom_y = 3 * dx
    else:
        nfy = round((2 * o + box[1]) / dx)
        parts.append(_lattice(nfx, nfy, LAYERS, 0.0, 0.0, 0.0, dx))         # floor
        parts.append(_lattice(LAYERS, nfy, wall_h, 0.0, 0.0, o, dx))        # left wall
        nbx = round(box[0] / dx)
        parts.append(_lattice(nbx, LAYERS, wall_h, o, 0.0, o, dx))          # side wall y = 0
        parts.append(_lattice(nbx, LAYERS, wall_h, o, o + box[1], o, dx))   # side wall y = max
        dom_y = 2 * o + box[1]
    walls = np.concatenate(parts)
    dp = dp_constants(phi, c, psi, E, nu)
    cs = math.sqrt(dp["K"] / rho0)
    k0 = 1.0 - math.sin(math.radians(phi))
    case = build_case(name, mat, walls, dx=dx, domain=(o + box[0] + dx, dom_y, o + z_margin * col[2]),
                      rho0=rho0, model=1, mpar=(dp["K"], dp["G"], dp["kphi"], dp["kpsi"], dp["kc"]), cs=cs,
                      shepard=shepard, nsteps=nsteps, z_top=o + col[2], k0=k0, alpha=alpha)
    case.meta.update(dict(dx=dx, col=col, box=box, quasi2d=quasi2d, phi=phi, c=c, floor_z=o, wall_x=o))
    return case


def column_2d_small(nsteps: int = 20000, seed: int = 0) -> Case:
    """Plane-strain sand column: 0.1 m x 0.2 m at dx = 2 mm (5,000 particles),
    +-1% jitter, 0.5 m box, phi = 19.8 deg, E = 0.84 MPa, nu = 0.3, rho0 = 2650,
    CFL 0.1, FP64 oracle.  The kernel stays 3-D (one y layer)."""
    return column(name="column_2d_small", dx=0.002, col=(0.1, 0.002, 0.2), box=(0.5, 0.002), quasi2d=True,
                  jitter=0.01, seed=seed, rho0=2650.0, E=0.84e6, nu=0.3, phi=19.8, c=0.0, psi=0.0,
                  shepard=0, nsteps=nsteps)


def column_2d_slab(nsteps: int = 20000, seed: int = 0, ny: int = 8) -> Case:
    """The plane-strain column of column_2d_small as a slab ny particles thick
    between two y side walls (the reference's plane example is a thick slab
    too: examples/sand_collapse_plane.xml holds 46 particles across y).  A
    one-particle slab under the always-3-D kernel is not supported by its
    floor; this slab is, so it carries the full-run comparison."""
    dx = 0.002
    return column(name="column_2d_slab", dx=dx, col=(0.1, ny * dx, 0.2), box=(0.5, ny * dx), quasi2d=False,
                  jitter=0.01, seed=seed, rho0=2650.0, E=0.84e6, nu=0.3, phi=19.8, c=0.0, psi=0.0,
                  shepard=0, nsteps=nsteps)


def column_3d(nsteps: int = 5000, dx: float = 0.005) -> Case:
    """3D column 0.4 x 0.4 x 0.8 m at dx = 5 mm (80 x 80 x 160 = 1,024,000
    material particles) in a 2.0 x 0.4 m box; floor, x = 0 and both side walls;
    phi = 30 deg, c = 0, psi = 0, E = 5 MPa, nu = 0.3; Shepard every 20 steps."""
    return column(name="column_3d", dx=dx, col=(0.4, 0.4, 0.8), box=(2.0, 0.4), quasi2d=False, rho0=2650.0,
                  E=5e6, nu=0.3, phi=30.0, c=0.0, psi=0.0, shepard=20, nsteps=nsteps, z_margin=1.15)


def small_column_3d(nx: int = 16, ny: int = 8, nz: int = 24, dx: float = 0.005, nsteps: int = 10,
                    jitter: float = 0.05, seed: int = 1, shepard: int = 0, phi: float = 30.0, c: float = 0.0,
                    box_x: float | None = None) -> Case:
    """Small 3D column for parity tests."""
    bx = box_x if box_x is not None else 3 * nx * dx
    return column(name="small_column_3d", dx=dx, col=(nx * dx, ny * dx, nz * dx), box=(bx, ny * dx),
                  quasi2d=False, jitter=jitter, seed=seed, shepard=shepard, nsteps=nsteps, phi=phi, c=c)


VW_SAND = dict(phi_c=33.0, hs=2.6e9, n=0.27, ed0=0.61, ec0=0.98, ei0=1.10, alpha=0.18, beta=1.0)


def slope_vw(npm_target: float = 4.0e6, nsteps: int = 10000, seed: int = 0, depth: float = 0.5,
             height: float = 1.0, length: float = 3.0, angle: float = 30.0, e0: float = 0.7,
             e_spread: float = 0.05, rho0: float = 2000.0, cs: float = 60.0) -> Case:
    """30-degree embankment (1.0 m high, 3.0 m base, 0.5 m deep), von
    Wolffersdorff hypoplasticity, per-particle void ratio e0 (1 +- 5%)."""
    run = height / math.tan(math.radians(angle))
    vol = 0.5 * (length + (length - run)) * height * depth
    dx = (vol / npm_target) ** (1.0 / 3.0)
    o = LAYERS * dx
    nx, ny, nz = math.ceil(length / dx), round(depth / dx), math.ceil(height / dx)
    pts = _lattice(nx, ny, nz, o, o, o, dx)
    rel_x, rel_z = pts[:, 0] - o, pts[:, 2] - o
    keep = (rel_z <= height) & (rel_x <= length - rel_z / math.tan(math.radians(angle)))
    mat = pts[keep]
    rng = np.random.default_rng(seed)
    ev = (e0 * (1.0 + rng.uniform(-e_spread, e_spread, len(mat)))).astype(np.float32)
    box_x = length + 1.5
    nfx, nfy = round((o + box_x) / dx), round((2 * o + depth) / dx)
    wall_h = round(height / dx)
    nbx = round(box_x / dx)
    walls = np.concatenate([
        _lattice(nfx, nfy, LAYERS, 0.0, 0.0, 0.0, dx),
        _lattice(LAYERS, nfy, wall_h, 0.0, 0.0, o, dx),
        _lattice(nbx, LAYERS, wall_h, o, 0.0, o, dx),
        _lattice(nbx, LAYERS, wall_h, o, o + depth, o, dx),
    ])
    # geostatic stress under the local surface: the crest, or the slope face
    # above the particle
    z_surf = o + np.minimum(height, (length - (mat[:, 0] - o)) * math.tan(math.radians(angle)))
    vw = abi.LqVw(**VW_SAND)
    case = build_case("slope_vw", mat, walls, dx=dx, domain=(o + box_x + dx, 2 * o + depth, o + 1.3 * height),
                      rho0=rho0, model=3, mpar=(0, 0, 0, 0, 0), cs=cs, shepard=0, nsteps=nsteps,
                      z_top=z_surf, k0=0.5, vw=vw, e0=ev)
    case.meta.update(dict(dx=dx, floor_z=o, wall_x=o))
    return case


def ensemble(nruns: int = 256, seed: int = 0, dx: float = 0.005, nsteps: int = 5000, npm_lo: float = 2e3,
             npm_hi: float = 6e4) -> list[Case]:
    """Independent 3D columns: npm log-uniform in [2k, 60k], aspect (height /
    width) U[0.5, 4], phi U[20, 40] deg, c U[0, 5] kPa."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nruns):
        npm = math.exp(rng.uniform(math.log(npm_lo), math.log(npm_hi)))
        aspect = rng.uniform(0.5, 4.0)
        phi = rng.uniform(20.0, 40.0)
        c = rng.uniform(0.0, 5e3)
        nw = max(2, round((npm / aspect) ** (1.0 / 3.0)))
        nh = max(2, round(aspect * nw))
        w, hgt = nw * dx, nh * dx
        case = column(name=f"ens{r}", dx=dx, col=(w, w, hgt), box=(w + 2.0 * hgt, w), quasi2d=False,
                      phi=phi, c=c, shepard=20, nsteps=nsteps, seed=seed * 1000 + r)
        out.append(case)
    return out


def served_mix(nreq: int = 48, step_scale: float = 1.0, seed: int = 0, rate: float = 20.0,
               extremes: bool = False) -> list[tuple[float, Case, int]]:
    """Request stream of the served workload: Poisson arrivals (rate per
    second), column (Drucker-Prager, phi U[20,40] deg, c U[0,5] kPa, aspect
    U[0.5,4]) or the 30-degree embankment of the slope workload (1.0 m high,
    3.0 m long, 0.5 m deep; von Wolffersdorff) 50/50, material
    particles log-normal around 30k clipped to [5k, 500k], 2,000-8,000 steps
    scaled by step_scale.  extremes: the first two requests take the clip
    bounds (5k and 500k particles) so that a short stream covers both ends.
    Returns (arrival time s, case, steps) in arrival order."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    out = []
    for i in range(nreq):
        npm = float(np.clip(math.exp(rng.normal(math.log(3e4), 1.0)), 5e3, 5e5))
        if extremes and i < 2:
            npm = (5e3, 5e5)[i]
        steps = max(1, int(round(rng.integers(2000, 8001) * step_scale)))
        if rng.random() < 0.5:
            aspect = rng.uniform(0.5, 4.0)
            nw = max(2, round((npm / aspect) ** (1 / 3)))
            w, hh = nw * 0.005, max(2, round(aspect * nw)) * 0.005
            c = column(name=f"req{i}", dx=0.005, col=(w, w, hh), box=(w + 2 * hh, w), quasi2d=False,
                       phi=rng.uniform(20, 40), c=rng.uniform(0, 5e3), shepard=20, nsteps=steps, seed=i)
        else:
            c = slope_vw(npm_target=npm, nsteps=steps, seed=i)
        out.append((float(arrivals[i]), c, steps))
    return out


def perturb(case: Case, seed: int = 7, v_amp: float = 0.2, s_amp: float = 2e3, rho_amp: float = 0.003) -> Case:
    """Random velocity, stress and density perturbation of the material
    particles, so that per-call parity tests exercise every pair term."""
    rng = np.random.default_rng(seed)
    npb, n = case.par.npb, case.par.np
    m = n - npb
    case.velrhop[npb:, :3] += rng.normal(0.0, v_amp, (m, 3)).astype(np.float32)
    case.velrhop[npb:, 3] *= (1.0 + rng.normal(0.0, rho_amp, m)).astype(np.float32)
    case.str[npb:] += rng.normal(0.0, s_amp, (m, 6)).astype(np.float32)
    case.stn[npb:] += rng.normal(0.0, 1e-4, (m, 6)).astype(np.float32)
    return case


# ------------------------------------------------------------ project files
_XML_KEYS_COMMON = [
    ("number-of-particles", "np", "int"), ("number-of-boundary-particles", "npb", "int"),
    ("number-of-material-particles", "npm", "int"),
    ("domain-size-x", "domain_size_x", "float"), ("domain-size-y", "domain_size_y", "float"),
    ("domain-size-z", "domain_size_z", "float"), ("domain-min-x", "domain_min_x", "float"),
    ("domain-min-y", "domain_min_y", "float"), ("domain-min-z", "domain_min_z", "float"),
    ("particle-size", "dr", "float"), ("h-over-dr", "hdr", "float"),
    ("reference-density", "rho0", "float"), ("speed-of-sound", "cs", "float"), ("cfl-number", "cfl", "float"),
    ("constitutive-model", "constitutive_model", "int"),
]
_XML_KEYS_DP = [("elastic-K", "model_par_1"), ("elastic-G", "model_par_2"), ("DP-kphi", "model_par_3"),
                ("DP-kpsi", "model_par_4"), ("DP-kcohesion", "model_par_5")]
_XML_KEYS_WU = [("hypo-par-c1", "model_par_1"), ("hypo-par-c2", "model_par_2"), ("hypo-par-c3", "model_par_3"),
                ("hypo-par-c4", "model_par_4"), ("cohesion", "model_par_5")]
_XML_KEYS_TAIL = [
    ("acceleration-x", "acc_x", "float"), ("acceleration-y", "acc_y", "float"), ("acceleration-z", "acc_z", "float"),
    ("time-max", "time_max", "float"), ("time", "time", "float"), ("time-step", "dt", "float"),
    ("output-frequency", "output_frequency", "float"), ("output-number", "output_number", "int"),
    ("density-shepard-filter", "density_shepard_filter", "int"),
    ("stress-shepard-filter", "stress_shepard_filter", "int"),
    ("dynamic-time-step", "dynamic_time_step", "int"), ("integration-method", "integration_method", "int"),
    ("viscosity-type", "viscosity_type", "int"),
]


def _fmt(v, t: str) -> str:
    return str(int(v)) if t == "int" else repr(float(np.float32(v)))


def write_xml(path: str, par: abi.LqParams, vw: abi.LqVw | None = None) -> None:
    lines = ['<?xml version="1.0" encoding="utf-8" ?>', "<parameters>"]
    add = lambda k, v, t: lines.append(f'  <{k} type="{t}">{_fmt(v, t)}</{k}>')
    for k, f, t in _XML_KEYS_COMMON:
        add(k, getattr(par, f), t)
    if par.constitutive_model == 1:
        for k, f in _XML_KEYS_DP:
            add(k, getattr(par, f), "float")
    elif par.constitutive_model == 2:
        for k, f in _XML_KEYS_WU:
            add(k, getattr(par, f), "float")
    elif par.constitutive_model == 3 and vw is not None:
        for k, f in (("vw-phi-c", "phi_c"), ("vw-hs", "hs"), ("vw-n", "n"), ("vw-ed0", "ed0"), ("vw-ec0", "ec0"),
                     ("vw-ei0", "ei0"), ("vw-alpha", "alpha"), ("vw-beta", "beta")):
            add(k, getattr(vw, f), "float")
    for k, f, t in _XML_KEYS_TAIL:
        add(k, getattr(par, f), t)
    if par.viscosity_type == 1:
        add("artificial-viscosity-alpha", par.artificial_viscosity_alpha, "float")
    add("artificial-pressure", par.artificial_pressure, "int")
    if par.artificial_pressure == 1:
        add("artificial-pressure-coefficient", par.artificial_pressure_coefficient, "float")
    add("output-format", par.output_format, "int")
    lines.append("</parameters>")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_dat(path: str, case: Case) -> None:
    cols = np.concatenate([case.pospres, case.velrhop, case.idep[:, :2], case.str, case.stn], 1)
    np.savetxt(path, cols, fmt="%.9g")


def write_project(prefix: str, case: Case) -> None:
    write_xml(prefix + ".xml", case.par, case.vw)
    write_dat(prefix + ".dat", case)
