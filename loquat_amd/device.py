"""Device arrays (torch tensors on the GPU) laid out as the C ABI expects,
and a thin wrapper of the engine.  torch is used only for device memory and
streams; every computation runs in the HIP library."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi
from .cases import Case


class DeviceState:
    """lq_state: FP64 positions + FP32 float4/tensor2d arrays + orig index."""

    def __init__(self, n: int, device: str = "cuda"):
        f32, f64 = torch.float32, torch.float64
        self.n = n
        self.xd = torch.zeros(n, dtype=f64, device=device)
        self.yd = torch.zeros(n, dtype=f64, device=device)
        self.zd = torch.zeros(n, dtype=f64, device=device)
        self.pospres = torch.zeros(n, 4, dtype=f32, device=device)
        self.velrhop = torch.zeros(n, 4, dtype=f32, device=device)
        self.idep = torch.zeros(n, 4, dtype=f32, device=device)
        self.str = torch.zeros(n, 6, dtype=f32, device=device)
        self.stn = torch.zeros(n, 6, dtype=f32, device=device)
        self.orig = torch.zeros(n, dtype=torch.int32, device=device)

    @classmethod
    def from_host(cls, posd, pospres, velrhop, idep, str_, stn, orig=None, device: str = "cuda") -> "DeviceState":
        n = len(posd)
        s = cls(n, device)
        pp = np.array(pospres, np.float32, copy=True)
        pp[:, :3] = posd.astype(np.float32)
        s.xd.copy_(torch.from_numpy(np.ascontiguousarray(posd[:, 0])))
        s.yd.copy_(torch.from_numpy(np.ascontiguousarray(posd[:, 1])))
        s.zd.copy_(torch.from_numpy(np.ascontiguousarray(posd[:, 2])))
        s.pospres.copy_(torch.from_numpy(pp))
        s.velrhop.copy_(torch.from_numpy(np.ascontiguousarray(velrhop, np.float32)))
        s.idep.copy_(torch.from_numpy(np.ascontiguousarray(idep, np.float32)))
        s.str.copy_(torch.from_numpy(np.ascontiguousarray(str_, np.float32)))
        s.stn.copy_(torch.from_numpy(np.ascontiguousarray(stn, np.float32)))
        o = np.arange(n, dtype=np.int32) if orig is None else np.asarray(orig, np.int32)
        s.orig.copy_(torch.from_numpy(o))
        return s

    @classmethod
    def from_case(cls, case: Case, device: str = "cuda") -> "DeviceState":
        return cls.from_host(case.posd, case.pospres, case.velrhop, case.idep, case.str, case.stn, device=device)

    def struct(self) -> abi.LqState:
        p = abi.ptr
        return abi.LqState(p(self.xd), p(self.yd), p(self.zd), p(self.pospres), p(self.velrhop), p(self.idep),
                           p(self.str), p(self.stn), p(self.orig))

    def host(self) -> dict:
        out = {k: getattr(self, k).cpu().numpy() for k in
               ("xd", "yd", "zd", "pospres", "velrhop", "idep", "str", "stn", "orig")}
        out["posd"] = np.stack([out["xd"], out["yd"], out["zd"]], 1)
        return out


class DeviceIni:
    def __init__(self, n: int, device: str = "cuda"):
        f32, f64 = torch.float32, torch.float64
        self.xd = torch.zeros(n, dtype=f64, device=device)
        self.yd = torch.zeros(n, dtype=f64, device=device)
        self.zd = torch.zeros(n, dtype=f64, device=device)
        self.velrhop = torch.zeros(n, 4, dtype=f32, device=device)
        self.epvr = torch.zeros(n, 2, dtype=f32, device=device)
        self.str = torch.zeros(n, 6, dtype=f32, device=device)
        self.stn = torch.zeros(n, 6, dtype=f32, device=device)

    def struct(self) -> abi.LqIni:
        p = abi.ptr
        return abi.LqIni(p(self.xd), p(self.yd), p(self.zd), p(self.velrhop), p(self.epvr), p(self.str),
                         p(self.stn))

    def host(self) -> dict:
        return {k: getattr(self, k).cpu().numpy() for k in ("xd", "yd", "zd", "velrhop", "epvr", "str", "stn")}


class DeviceRates:
    def __init__(self, n: int, device: str = "cuda"):
        self.accad = torch.zeros(n, 4, dtype=torch.float32, device=device)
        self.spin = torch.zeros(n, 4, dtype=torch.float32, device=device)
        self.acestn = torch.zeros(n, 6, dtype=torch.float32, device=device)

    def struct(self) -> abi.LqRates:
        p = abi.ptr
        return abi.LqRates(p(self.accad), p(self.spin), p(self.acestn))

    def host(self) -> dict:
        return {k: getattr(self, k).cpu().numpy() for k in ("accad", "spin", "acestn")}


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


class Grid:
    """lq_grid handle (LoquatGrid)."""

    def __init__(self, np_cap: int, ncell_cap: int):
        self.h = C.c_void_p()
        abi.check(abi.lib().lq_grid_create(np_cap, ncell_cap, C.byref(self.h)))

    def build(self, par: abi.LqParams, st: DeviceState) -> None:
        abi.check(abi.lib().lq_grid_build(self.h, C.byref(par), abi.ptr(st.pospres), stream_ptr()))

    def permute(self, n: int, src: DeviceState, dst: DeviceState, ini_src: DeviceIni | None = None,
                ini_dst: DeviceIni | None = None) -> None:
        a, b = src.struct(), dst.struct()
        ia = C.byref(ini_src.struct()) if ini_src else None
        ib = C.byref(ini_dst.struct()) if ini_dst else None
        abi.check(abi.lib().lq_grid_permute(self.h, n, C.byref(a), C.byref(b), ia, ib, stream_ptr()))

    def arrays(self, n: int, ncell: int) -> dict:
        """Copies part_cel, part_idx, cell_beg, cell_end to host numpy."""
        ptrs = [C.c_void_p() for _ in range(4)]
        abi.check(abi.lib().lq_grid_arrays(self.h, *[C.byref(p) for p in ptrs]))
        torch.cuda.synchronize()
        out = {}
        for name, p, cnt in zip(("part_cel", "part_idx", "cell_beg", "cell_end"), ptrs, (n, n, ncell, ncell)):
            buf = np.empty(cnt, np.int32)
            if cnt:
                _memcpy_d2h(buf, p.value)
            out[name] = buf
        return out

    def table_ptrs(self) -> tuple[int, int]:
        ptrs = [C.c_void_p() for _ in range(4)]
        abi.check(abi.lib().lq_grid_arrays(self.h, *[C.byref(p) for p in ptrs]))
        return ptrs[2].value, ptrs[3].value

    def close(self) -> None:
        if self.h:
            abi.lib().lq_grid_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _memcpy_d2h(dst: np.ndarray, src_ptr: int) -> None:
    """Device -> host copy through torch (wraps the raw device pointer)."""
    n = dst.size
    # torch has no from-pointer constructor; copy via a hip memcpy through ctypes on libamdhip64.
    _hip().hipMemcpy(C.c_void_p(dst.ctypes.data), C.c_void_p(src_ptr), C.c_size_t(n * dst.itemsize), 2)


_hiplib = None


def _hip():
    global _hiplib
    if _hiplib is None:
        _hiplib = C.CDLL("libamdhip64.so")
        _hiplib.hipMemcpy.restype = C.c_int
        _hiplib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hiplib


# ------------------------------------------------------------ module calls
def adami(par, st: DeviceState, beg: int, end: int) -> None:
    s = st.struct()
    abi.check(abi.lib().lq_adami_boundary(C.byref(par), C.byref(s), beg, end, stream_ptr()))


def interaction(par, st: DeviceState, rt: DeviceRates, beg: int, end: int, nbr: torch.Tensor | None = None) -> None:
    s, r = st.struct(), rt.struct()
    abi.check(abi.lib().lq_particle_interaction(C.byref(par), C.byref(s), C.byref(r), beg, end, abi.ptr(nbr),
                                                stream_ptr()))


def interaction_stale(par, st: DeviceState, rt: DeviceRates, then: tuple, disp: float, beg: int, end: int,
                      nbr: torch.Tensor | None = None) -> None:
    """lq_particle_interaction_stale: then = (x, y, z) FP64 device tensors of
    the positions the table was built from."""
    s, r = st.struct(), rt.struct()
    xo, yo, zo = (abi.ptr(t) for t in then)
    abi.check(abi.lib().lq_particle_interaction_stale(C.byref(par), C.byref(s), C.byref(r), xo, yo, zo,
                                                      C.c_float(disp), beg, end, abi.ptr(nbr), stream_ptr()))


def adami_stale(par, st: DeviceState, disp: float, beg: int, end: int, nbr: torch.Tensor | None = None) -> None:
    s = st.struct()
    abi.check(abi.lib().lq_adami_boundary_stale(C.byref(par), C.byref(s), C.c_float(disp), beg, end,
                                                abi.ptr(nbr), stream_ptr()))


def predictor(par, vw, st: DeviceState, ini: DeviceIni, rt: DeviceRates) -> None:
    s, i, r = st.struct(), ini.struct(), rt.struct()
    abi.check(abi.lib().lq_predictor(C.byref(par), C.byref(vw), C.byref(s), C.byref(i), C.byref(r), stream_ptr()))


def corrector(par, vw, st: DeviceState, ini: DeviceIni, rt: DeviceRates) -> None:
    s, i, r = st.struct(), ini.struct(), rt.struct()
    abi.check(abi.lib().lq_corrector(C.byref(par), C.byref(vw), C.byref(s), C.byref(i), C.byref(r), stream_ptr()))


def smoothing(par, st: DeviceState, cache: torch.Tensor, beg: int, end: int) -> None:
    s = st.struct()
    abi.check(abi.lib().lq_field_smoothing(C.byref(par), C.byref(s), abi.ptr(cache), beg, end, stream_ptr()))


class Engine:
    """lq_engine: one or many runs, stepped together."""

    def __init__(self, np_cap: int, ncell_cap: int, max_runs: int = 1):
        self.h = C.c_void_p()
        abi.check(abi.lib().lq_engine_create(int(np_cap), int(ncell_cap), int(max_runs), C.byref(self.h)))
        self._keep = {}

    def add(self, case: Case, nsteps: int | None = None) -> int:
        ri = case.run_input(nsteps)
        rid = C.c_int32()
        abi.check(abi.lib().lq_engine_add_run(self.h, C.byref(ri), C.byref(rid)))
        return rid.value

    def remove(self, rid: int) -> None:
        abi.check(abi.lib().lq_engine_remove_run(self.h, rid))

    def step(self, n: int = 1) -> None:
        abi.check(abi.lib().lq_engine_step(self.h, int(n)))

    def sync(self) -> None:
        abi.check(abi.lib().lq_engine_synchronize(self.h))

    def set_graphs(self, on: bool) -> None:
        abi.check(abi.lib().lq_engine_set_graphs(self.h, 1 if on else 0))

    def compact(self) -> int:
        """Removes material particles that left their run's grid; returns how many."""
        n = C.c_int64()
        abi.check(abi.lib().lq_engine_compact(self.h, C.byref(n)))
        return n.value

    def poll(self) -> int:
        """Collects the device's failure flags; returns the failed resident runs."""
        n = C.c_int32()
        abi.check(abi.lib().lq_engine_poll(self.h, C.byref(n)))
        return n.value

    def failed(self, rid: int) -> bool:
        f = C.c_int32()
        abi.check(abi.lib().lq_engine_run_failed(self.h, rid, C.byref(f)))
        return bool(f.value)

    def set_variant(self, variant: int) -> None:
        """Interaction variant of this engine (0 .. abi.INTERACTION_VARIANTS - 1)."""
        abi.check(abi.lib().lq_engine_set_interaction_variant(self.h, int(variant)))

    def variant(self) -> int:
        return abi.lib().lq_engine_get_interaction_variant(self.h)

    def set_cell_order(self, zbin_bits: int) -> None:
        """0: the reference's in-cell order; b > 0: cells sorted by 2^b z-bins."""
        abi.check(abi.lib().lq_engine_set_cell_order(self.h, int(zbin_bits)))

    def neighbour_counts(self, rid: int) -> np.ndarray:
        """Interaction neighbour count per particle of run rid (input order)."""
        out = np.zeros(self.info(rid)["np"], np.int32)
        abi.check(abi.lib().lq_engine_neighbour_counts(self.h, rid, out.ctypes.data))
        return out

    def neighbour_counts_stale(self, rid: int) -> np.ndarray:
        """The same counts from the cell tables the runs keep between sorts
        (no sort; as the next step's predictor half scans them)."""
        out = np.zeros(self.info(rid)["np"], np.int32)
        abi.check(abi.lib().lq_engine_neighbour_counts_stale(self.h, rid, out.ctypes.data))
        return out

    def set_column_order(self, strip_rows: int) -> None:
        """(x, y) column order of the engine's sort: strips of strip_rows x-rows
        (0 or 1: cell id order).  A run gives the same bits under every order;
        the re-sort this call makes mid-run changes the order of FP32 sums."""
        abi.check(abi.lib().lq_engine_set_column_order(self.h, int(strip_rows)))

    def column_order(self) -> int:
        return abi.lib().lq_engine_get_column_order(self.h)

    def set_sort_interval(self, steps: int) -> None:
        """Sort each run (and rebuild its cell table) every `steps` of its own steps."""
        abi.check(abi.lib().lq_engine_set_sort_interval(self.h, int(steps)))

    def sort_interval(self) -> int:
        return abi.lib().lq_engine_get_sort_interval(self.h)

    def info(self, rid: int) -> dict:
        s, t, f, n = C.c_int64(), C.c_double(), C.c_int32(), C.c_int32()
        abi.check(abi.lib().lq_engine_run_info(self.h, rid, C.byref(s), C.byref(t), C.byref(f), C.byref(n)))
        return dict(steps=s.value, time=t.value, finished=bool(f.value), np=n.value)

    def totals(self) -> dict:
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        abi.check(abi.lib().lq_engine_totals(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(npm=a.value, np=b.value, ncell=c.value, nruns=d.value)

    def download(self, rid: int) -> dict:
        n = self.info(rid)["np"]
        out = {k: np.zeros((n, w), np.float32) for k, w in
               (("pospres", 4), ("velrhop", 4), ("idep", 4), ("str", 6), ("stn", 6))}
        posd = np.zeros((n, 3), np.float64)
        hp = abi.host_particles(out["pospres"], out["velrhop"], out["idep"], out["str"], out["stn"])
        abi.check(abi.lib().lq_engine_download(self.h, rid, C.byref(hp), posd.ctypes.data))
        out["posd"] = posd
        return out

    def autotune(self, nrep: int = 3) -> tuple[int, list[float]]:
        """Pins the fastest interaction variant for this engine."""
        best = C.c_int32()
        ms = (C.c_double * abi.INTERACTION_VARIANTS)()
        abi.check(abi.lib().lq_engine_autotune(self.h, int(nrep), C.byref(best), ms))
        return best.value, list(ms)

    def autotune_cell_order(self, nrep: int = 3) -> tuple[int, list[float]]:
        """Pins the fastest in-cell order (2^b height bins per cell) for this
        engine, timing sort, permute, Adami and interaction per candidate b."""
        best = C.c_int32()
        ms = (C.c_double * abi.CELL_ORDER_CANDIDATES)()
        abi.check(abi.lib().lq_engine_autotune_cell_order(self.h, int(nrep), C.byref(best), ms))
        return best.value, list(ms)

    def autotune_column_order(self, nrep: int = 3) -> tuple[int, list[float]]:
        """Pins the fastest (x, y) column order (strips of 0, 2, 4, 8, 16
        rows) for this engine, timing sort, Adami and interaction per
        candidate; results do not depend on it."""
        best = C.c_int32()
        ms = (C.c_double * abi.COLUMN_ORDER_CANDIDATES)()
        abi.check(abi.lib().lq_engine_autotune_column_order(self.h, int(nrep), C.byref(best), ms))
        return best.value, list(ms)

    def profile(self, nsteps: int) -> list[float]:
        ms = (C.c_double * 5)()
        abi.check(abi.lib().lq_engine_profile(self.h, int(nsteps), ms))
        return list(ms)

    def stream(self) -> int:
        return abi.lib().lq_engine_stream(self.h)

    def close(self) -> None:
        if self.h:
            abi.lib().lq_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bytes_per_run(np_: int, ncell: int) -> int:
    return abi.lib().lq_engine_bytes_per_run(int(np_), int(ncell))
