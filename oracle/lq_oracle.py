"""ctypes front end of the FP64 CPU oracle (oracle/lq_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblq_oracle.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)


class _State(C.Structure):
    _fields_ = [(k, dp) for k in ("x", "y", "z", "p", "vx", "vy", "vz", "rho", "id", "ep", "voidr")] + \
               [("s", dp * 6), ("e", dp * 6), ("orig", ip)]


class _Ini(C.Structure):
    _fields_ = [(k, dp) for k in ("x", "y", "z", "vx", "vy", "vz", "rho", "ep", "voidr")] + \
               [("s", dp * 6), ("e", dp * 6)]


class _Rates(C.Structure):
    _fields_ = [(k, dp) for k in ("ax", "ay", "az", "drho", "wx", "wy", "wz")] + [("d", dp * 6)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liblq_oracle.so"], check=True)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.lqo_cell_keys.argtypes = [vp, C.c_int, dp, dp, dp, ip]
        L.lqo_stable_sort.argtypes = [C.c_int, ip, C.c_int, ip, ip]
        L.lqo_cell_table.argtypes = [C.c_int, ip, C.c_int, ip, ip]
        L.lqo_grid_build.argtypes = [vp, C.c_int, dp, dp, dp, ip, ip, ip, ip]
        L.lqo_permute.argtypes = [C.c_int, ip, C.POINTER(_State), C.POINTER(_Ini)]
        L.lqo_adami.argtypes = [vp, C.POINTER(_State), ip, ip]
        L.lqo_interaction.argtypes = [vp, C.POINTER(_State), C.POINTER(_Rates), ip, ip, ip]
        L.lqo_interaction_abs.argtypes = [vp, C.POINTER(_State), C.POINTER(_Rates), ip, ip]
        L.lqo_borderline_pairs.argtypes = [vp, C.POINTER(_State), ip, ip, C.POINTER(C.c_int64)]
        L.lqo_predictor.argtypes = [vp, vp, C.POINTER(_State), C.POINTER(_Ini), C.POINTER(_Rates)]
        L.lqo_corrector.argtypes = [vp, vp, C.POINTER(_State), C.POINTER(_Ini), C.POINTER(_Rates)]
        L.lqo_shepard.argtypes = [vp, C.POINTER(_State), ip, ip]
        L.lqo_initial_sort.argtypes = [vp, C.POINTER(_State)]
        L.lqo_run.argtypes = [vp, vp, C.POINTER(_State), C.c_int, C.c_int]
        L.lqo_run_dt.argtypes = [vp, vp, C.POINTER(_State), C.c_int, C.c_int, dp]
        L.lqo_dynamic_dt.argtypes = [vp, C.POINTER(_State), C.POINTER(_Rates), C.c_int]
        L.lqo_dynamic_dt.restype = C.c_double
        L.lqo_num_threads.restype = C.c_int
        L.lqo_set_num_threads.argtypes = [C.c_int]
        L.lqo_constitutive.argtypes = [vp, vp, C.c_double, dp, dp, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ip)


class OState:
    """FP64 state, SoA."""

    SCALARS = ("x", "y", "z", "p", "vx", "vy", "vz", "rho", "id", "ep", "voidr")

    def __init__(self, n: int):
        self.n = n
        for k in self.SCALARS:
            setattr(self, k, np.zeros(n, np.float64))
        self.s = np.zeros((6, n), np.float64)
        self.e = np.zeros((6, n), np.float64)
        self.orig = np.arange(n, dtype=np.int32)

    @classmethod
    def from_arrays(cls, posd, pospres, velrhop, idep, str_, stn, orig=None) -> "OState":
        n = len(posd)
        o = cls(n)
        posd = np.asarray(posd, np.float64)
        o.x[:], o.y[:], o.z[:] = posd[:, 0], posd[:, 1], posd[:, 2]
        o.p[:] = pospres[:, 3]
        o.vx[:], o.vy[:], o.vz[:], o.rho[:] = velrhop[:, 0], velrhop[:, 1], velrhop[:, 2], velrhop[:, 3]
        o.id[:], o.ep[:], o.voidr[:] = idep[:, 0], idep[:, 1], idep[:, 2]
        o.s[:] = np.asarray(str_, np.float64).T
        o.e[:] = np.asarray(stn, np.float64).T
        if orig is not None:
            o.orig[:] = orig
        return o

    @classmethod
    def from_case(cls, case) -> "OState":
        return cls.from_arrays(case.posd, case.pospres, case.velrhop, case.idep, case.str, case.stn)

    @classmethod
    def from_device(cls, h: dict) -> "OState":
        return cls.from_arrays(h["posd"], h["pospres"], h["velrhop"], h["idep"], h["str"], h["stn"], h["orig"])

    def struct(self) -> _State:
        st = _State(*[_d(getattr(self, k)) for k in self.SCALARS])
        for c in range(6):
            st.s[c] = _d(self.s[c])
            st.e[c] = _d(self.e[c])
        st.orig = _i(self.orig)
        return st

    def copy(self) -> "OState":
        o = OState(self.n)
        for k in self.SCALARS:
            getattr(o, k)[:] = getattr(self, k)
        o.s[:], o.e[:], o.orig[:] = self.s, self.e, self.orig
        return o

    @property
    def posd(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], 1)


class OIni:
    SCALARS = ("x", "y", "z", "vx", "vy", "vz", "rho", "ep", "voidr")

    def __init__(self, n: int):
        for k in self.SCALARS:
            setattr(self, k, np.zeros(n, np.float64))
        self.s = np.zeros((6, n), np.float64)
        self.e = np.zeros((6, n), np.float64)

    @classmethod
    def from_device(cls, h: dict) -> "OIni":
        n = len(h["xd"])
        o = cls(n)
        o.x[:], o.y[:], o.z[:] = h["xd"], h["yd"], h["zd"]
        v = h["velrhop"]
        o.vx[:], o.vy[:], o.vz[:], o.rho[:] = v[:, 0], v[:, 1], v[:, 2], v[:, 3]
        o.ep[:], o.voidr[:] = h["epvr"][:, 0], h["epvr"][:, 1]
        o.s[:] = h["str"].T
        o.e[:] = h["stn"].T
        return o

    def struct(self) -> _Ini:
        st = _Ini(*[_d(getattr(self, k)) for k in self.SCALARS])
        for c in range(6):
            st.s[c] = _d(self.s[c])
            st.e[c] = _d(self.e[c])
        return st


class ORates:
    SCALARS = ("ax", "ay", "az", "drho", "wx", "wy", "wz")

    def __init__(self, n: int):
        for k in self.SCALARS:
            setattr(self, k, np.zeros(n, np.float64))
        self.d = np.zeros((6, n), np.float64)

    @classmethod
    def from_device(cls, h: dict) -> "ORates":
        n = len(h["accad"])
        o = cls(n)
        a, w = h["accad"], h["spin"]
        o.ax[:], o.ay[:], o.az[:], o.drho[:] = a[:, 0], a[:, 1], a[:, 2], a[:, 3]
        o.wx[:], o.wy[:], o.wz[:] = w[:, 0], w[:, 1], w[:, 2]
        o.d[:] = h["acestn"].T
        return o

    def struct(self) -> _Rates:
        st = _Rates(*[_d(getattr(self, k)) for k in self.SCALARS])
        for c in range(6):
            st.d[c] = _d(self.d[c])
        return st


def _p(par):
    if par is None:
        return None
    return C.cast(C.pointer(par), C.c_void_p)


def cell_keys(par, st: OState) -> np.ndarray:
    k = np.zeros(st.n, np.int32)
    lib().lqo_cell_keys(_p(par), st.n, _d(st.x), _d(st.y), _d(st.z), _i(k))
    return k


def stable_sort(keys: np.ndarray, ncell: int) -> tuple[np.ndarray, np.ndarray]:
    n = len(keys)
    perm, srt = np.zeros(n, np.int32), np.zeros(n, np.int32)
    lib().lqo_stable_sort(n, _i(np.ascontiguousarray(keys, np.int32)), ncell, _i(perm), _i(srt))
    return perm, srt


def cell_table(sorted_keys: np.ndarray, ncell: int) -> tuple[np.ndarray, np.ndarray]:
    beg, end = np.zeros(ncell, np.int32), np.zeros(ncell, np.int32)
    lib().lqo_cell_table(len(sorted_keys), _i(np.ascontiguousarray(sorted_keys, np.int32)), ncell, _i(beg), _i(end))
    return beg, end


def grid_build(par, st: OState) -> dict:
    n, nc = st.n, par.ncell
    perm, srt = np.zeros(n, np.int32), np.zeros(n, np.int32)
    beg, end = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
    lib().lqo_grid_build(_p(par), n, _d(st.x), _d(st.y), _d(st.z), _i(perm), _i(srt), _i(beg), _i(end))
    return dict(part_idx=perm, part_cel=srt, cell_beg=beg, cell_end=end)


def permute(perm: np.ndarray, st: OState, ini: OIni | None = None) -> None:
    s = st.struct()
    i = C.byref(ini.struct()) if ini is not None else None
    lib().lqo_permute(st.n, _i(np.ascontiguousarray(perm, np.int32)), C.byref(s), i)


def adami(par, st: OState, beg, end) -> None:
    s = st.struct()
    lib().lqo_adami(_p(par), C.byref(s), _i(beg), _i(end))


def interaction(par, st: OState, beg, end) -> tuple[ORates, np.ndarray]:
    rt = ORates(st.n)
    nbr = np.zeros(st.n, np.int32)
    s, r = st.struct(), rt.struct()
    lib().lqo_interaction(_p(par), C.byref(s), C.byref(r), _i(beg), _i(end), _i(nbr))
    return rt, nbr


def interaction_abs(par, st: OState, beg, end) -> ORates:
    """Per material particle: sum over pairs of |term| for ax, ay, az (plus
    |g|) and drho -- the scale of a cancelling sum's rounding error."""
    rt = ORates(st.n)
    s, r = st.struct(), rt.struct()
    lib().lqo_interaction_abs(_p(par), C.byref(s), C.byref(r), _i(beg), _i(end))
    return rt


def borderline_pairs(par, st: OState, beg, end) -> tuple[int, int]:
    """(pairs the FP32 predicate accepts and FP64 rejects, the reverse),
    ordered pairs over the material particles' stencils."""
    out = (C.c_int64 * 2)()
    s = st.struct()
    lib().lqo_borderline_pairs(_p(par), C.byref(s), _i(beg), _i(end), out)
    return int(out[0]), int(out[1])


def predictor(par, vw, st: OState, ini: OIni, rt: ORates) -> None:
    s, i, r = st.struct(), ini.struct(), rt.struct()
    lib().lqo_predictor(_p(par), _p(vw), C.byref(s), C.byref(i), C.byref(r))


def corrector(par, vw, st: OState, ini: OIni, rt: ORates) -> None:
    s, i, r = st.struct(), ini.struct(), rt.struct()
    lib().lqo_corrector(_p(par), _p(vw), C.byref(s), C.byref(i), C.byref(r))


def shepard(par, st: OState, beg, end) -> None:
    s = st.struct()
    lib().lqo_shepard(_p(par), C.byref(s), _i(beg), _i(end))


def initial_sort(par, st: OState) -> None:
    s = st.struct()
    lib().lqo_initial_sort(_p(par), C.byref(s))


def run(par, vw, st: OState, nsteps: int, step0: int = 0) -> None:
    s = st.struct()
    lib().lqo_run(_p(par), _p(vw), C.byref(s), int(nsteps), int(step0))


def run_dt(par, vw, st: OState, nsteps: int, step0: int = 0) -> np.ndarray:
    """run(), returning the dt of every step (varies with dynamic-time-step = 1)."""
    dts = np.zeros(max(nsteps, 1), np.float64)
    s = st.struct()
    lib().lqo_run_dt(_p(par), _p(vw), C.byref(s), int(nsteps), int(step0), _d(dts))
    return dts[:nsteps]


def dynamic_dt(par, st: OState, rt: ORates, use_rates: bool) -> float:
    s, r = st.struct(), rt.struct()
    return lib().lqo_dynamic_dt(_p(par), C.byref(s), C.byref(r), 1 if use_rates else 0)


def num_threads() -> int:
    return lib().lqo_num_threads()


def set_num_threads(n: int) -> None:
    lib().lqo_set_num_threads(int(n))


def unpermute(st: OState) -> OState:
    """State in input order (by orig)."""
    o = OState(st.n)
    idx = st.orig
    for k in OState.SCALARS:
        getattr(o, k)[idx] = getattr(st, k)
    o.s[:, idx] = st.s
    o.e[:, idx] = st.e
    o.orig[:] = np.arange(st.n, dtype=np.int32)
    return o


def constitutive(par, vw, dt: float, spin3, d6, s6, stn6=None, p: float = 0.0, ep: float = 0.0,
                 voidr: float = 0.0):
    """Single-point constitutive update; returns (stress[6], p, ep, voidr)."""
    w = np.ascontiguousarray(spin3, np.float64)
    d = np.ascontiguousarray(d6, np.float64)
    n = np.ascontiguousarray(np.zeros(6) if stn6 is None else stn6, np.float64)
    s = np.array(s6, np.float64)
    pp, ee, vv = C.c_double(p), C.c_double(ep), C.c_double(voidr)
    vwp = _p(vw) if vw is not None else None
    lib().lqo_constitutive(_p(par), vwp, float(dt), _d(w), _d(d), _d(n), C.byref(pp), C.byref(ee), C.byref(vv),
                           _d(s))
    return s, pp.value, ee.value, vv.value
