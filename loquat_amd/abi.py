"""ctypes bindings of the C ABI declared in include/loquat_amd.h.

The library must be built (python -m loquat_amd.build); loading fails loudly
otherwise -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libloquat_amd.so")

i32, f32, f64, vp = C.c_int32, C.c_float, C.c_double, C.c_void_p

PARAM_FIELDS = [
    ("np", i32), ("npb", i32), ("npm", i32), ("np_delete", i32),
    ("domain_size_x", f32), ("domain_size_y", f32), ("domain_size_z", f32),
    ("domain_min_x", f32), ("domain_min_y", f32), ("domain_min_z", f32),
    ("cell_num_x", i32), ("cell_num_y", i32), ("cell_num_z", i32),
    ("ncell", i32), ("ncell_one_layer", i32), ("ncell_delete", i32),
    ("h", f32), ("i_h", f32), ("eta", f32), ("dr", f32), ("hdr", f32),
    ("m", f32), ("cs", f32), ("cfl", f32), ("rho0", f32),
    ("constitutive_model", i32),
    ("model_par_1", f32), ("model_par_2", f32), ("model_par_3", f32), ("model_par_4", f32), ("model_par_5", f32),
    ("density_shepard_filter", i32), ("stress_shepard_filter", i32),
    ("dynamic_time_step", i32), ("integration_method", i32), ("viscosity_type", i32),
    ("artificial_viscosity_alpha", f32),
    ("artificial_pressure", i32),
    ("artificial_pressure_coefficient", f32),
    ("acc_x", f32), ("acc_y", f32), ("acc_z", f32),
    ("time_max", f32), ("time", f32), ("dt", f32), ("output_frequency", f32),
    ("output_number", i32),
    ("kernel_normalization_par", f32), ("kernel_gradient_normalization_par", f32),
    ("kernel_zero", f32), ("kernel_dr", f32),
    ("output_format", i32),
]


class LqParams(C.Structure):
    _fields_ = PARAM_FIELDS

    def copy(self) -> "LqParams":
        p = LqParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(LqParams))
        return p

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in PARAM_FIELDS}


assert C.sizeof(LqParams) == 208


class LqVw(C.Structure):
    _fields_ = [("phi_c", f32), ("hs", f32), ("n", f32), ("ed0", f32), ("ec0", f32), ("ei0", f32),
                ("alpha", f32), ("beta", f32)]


class LqState(C.Structure):
    _fields_ = [("xd", vp), ("yd", vp), ("zd", vp), ("pospres", vp), ("velrhop", vp), ("idep", vp),
                ("str", vp), ("stn", vp), ("orig", vp)]


class LqIni(C.Structure):
    _fields_ = [("xd", vp), ("yd", vp), ("zd", vp), ("velrhop", vp), ("epvr", vp), ("str", vp), ("stn", vp)]


class LqRates(C.Structure):
    _fields_ = [("accad", vp), ("spin", vp), ("acestn", vp)]


class LqHostParticles(C.Structure):
    _fields_ = [("pospres", vp), ("velrhop", vp), ("idep", vp), ("str", vp), ("stn", vp)]


class LqRunInput(C.Structure):
    _fields_ = [("par", LqParams), ("vw", LqVw), ("host", LqHostParticles), ("posd", vp), ("nsteps", C.c_int64)]


class LqServerConfig(C.Structure):
    _fields_ = [("np_cap", C.c_int64), ("ncell_cap", C.c_int64), ("max_runs", i32), ("hbm_budget", C.c_int64),
                ("steps_per_tick", i32), ("fake_backend", i32), ("keep_results", i32)]


REQ_STATE = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "rejected"}
INTERACTION_VARIANTS = 20  # LQ_INTERACTION_VARIANTS
CELL_ORDER_CANDIDATES = 6  # LQ_CELL_ORDER_CANDIDATES
DEFAULT_INTERACTION_VARIANT = 19  # LQ_DEFAULT_INTERACTION_VARIANT
COLUMN_ORDER_CANDIDATES = 5  # LQ_COLUMN_ORDER_CANDIDATES
DEFAULT_SORT_INTERVAL = 12  # LQ_DEFAULT_SORT_INTERVAL
MAX_SORT_INTERVAL = 16  # LQ_MAX_SORT_INTERVAL
MAX_PARTICLES = 178956970  # LQ_MAX_PARTICLES
NOALLOC_MAX_PARTICLES = 2097152  # LQ_NOALLOC_MAX_PARTICLES


class LqError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"lq_status {status}: {msg}")
        self.status = status


STATUS = {0: "OK", 1: "INVALID_ARG", 2: "HIP", 3: "OUT_OF_MEMORY", 4: "IO", 5: "PARSE", 6: "MISSING_KEY",
          7: "UNSUPPORTED", 8: "CAPACITY", 9: "DOMAIN", 10: "STATE"}

_lib = None

# name -> (restype, argtypes)
_PROTOS = {
    "lq_params_derive": (i32, [C.POINTER(LqParams)]),
    "lq_neighbour_r2_cut": (i32, [C.POINTER(LqParams), C.POINTER(f32)]),
    "lq_load_parameters": (i32, [C.c_char_p, C.POINTER(LqParams), C.POINTER(LqVw)]),
    "lq_load_particles": (i32, [C.c_char_p, C.POINTER(LqParams), C.POINTER(LqHostParticles)]),
    "lq_last_error": (C.c_char_p, []),
    "lq_save_output": (i32, [C.c_char_p, C.POINTER(LqParams), C.POINTER(LqVw), C.POINTER(LqHostParticles), vp]),
    "lq_load_restart": (i32, [C.c_char_p, C.POINTER(LqParams), vp, vp]),
    "lq_grid_create": (i32, [C.c_int64, C.c_int64, C.POINTER(vp)]),
    "lq_grid_destroy": (i32, [vp]),
    "lq_grid_build": (i32, [vp, C.POINTER(LqParams), vp, vp]),
    "lq_grid_permute": (i32, [vp, i32, C.POINTER(LqState), C.POINTER(LqState), C.POINTER(LqIni),
                              C.POINTER(LqIni), vp]),
    "lq_grid_arrays": (i32, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "lq_adami_boundary": (i32, [C.POINTER(LqParams), C.POINTER(LqState), vp, vp, vp]),
    "lq_adami_boundary_stale": (i32, [C.POINTER(LqParams), C.POINTER(LqState), C.c_float, vp, vp, vp, vp]),
    "lq_particle_interaction_stale": (i32, [C.POINTER(LqParams), C.POINTER(LqState), C.POINTER(LqRates), vp, vp,
                                            vp, C.c_float, vp, vp, vp, vp]),
    "lq_particle_interaction": (i32, [C.POINTER(LqParams), C.POINTER(LqState), C.POINTER(LqRates), vp, vp, vp,
                                      vp]),
    "lq_predictor": (i32, [C.POINTER(LqParams), C.POINTER(LqVw), C.POINTER(LqState), C.POINTER(LqIni),
                           C.POINTER(LqRates), vp]),
    "lq_corrector": (i32, [C.POINTER(LqParams), C.POINTER(LqVw), C.POINTER(LqState), C.POINTER(LqIni),
                           C.POINTER(LqRates), vp]),
    "lq_field_smoothing": (i32, [C.POINTER(LqParams), C.POINTER(LqState), vp, vp, vp, vp]),
    "lq_engine_create": (i32, [C.c_int64, C.c_int64, i32, C.POINTER(vp)]),
    "lq_engine_destroy": (i32, [vp]),
    "lq_engine_add_run": (i32, [vp, C.POINTER(LqRunInput), C.POINTER(i32)]),
    "lq_engine_remove_run": (i32, [vp, i32]),
    "lq_engine_step": (i32, [vp, i32]),
    "lq_engine_synchronize": (i32, [vp]),
    "lq_engine_run_info": (i32, [vp, i32, C.POINTER(C.c_int64), C.POINTER(f64), C.POINTER(i32), C.POINTER(i32)]),
    "lq_engine_download": (i32, [vp, i32, C.POINTER(LqHostParticles), vp]),
    "lq_engine_totals": (i32, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                               C.POINTER(i32)]),
    "lq_engine_bytes_per_run": (C.c_int64, [C.c_int64, C.c_int64]),
    "lq_engine_stream": (vp, [vp]),
    "lq_engine_set_graphs": (i32, [vp, i32]),
    "lq_engine_set_cell_order": (i32, [vp, i32]),
    "lq_engine_neighbour_counts": (i32, [vp, i32, vp]),
    "lq_engine_profile": (i32, [vp, i32, C.POINTER(f64)]),
    "lq_engine_poll": (i32, [vp, C.POINTER(i32)]),
    "lq_engine_compact": (i32, [vp, C.POINTER(C.c_int64)]),
    "lq_engine_autotune": (i32, [vp, i32, C.POINTER(i32), C.POINTER(f64)]),
    "lq_engine_autotune_cell_order": (i32, [vp, i32, C.POINTER(i32), C.POINTER(f64)]),
    "lq_engine_run_failed": (i32, [vp, i32, C.POINTER(i32)]),
    "lq_server_create": (i32, [C.POINTER(LqServerConfig), C.POINTER(vp)]),
    "lq_server_destroy": (i32, [vp]),
    "lq_server_submit": (i32, [vp, C.POINTER(LqRunInput), C.POINTER(i32)]),
    "lq_server_tick": (i32, [vp, C.POINTER(i32), i32, C.POINTER(i32)]),
    "lq_server_request_info": (i32, [vp, i32, C.POINTER(i32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]),
    "lq_server_result": (i32, [vp, i32, C.POINTER(LqHostParticles), vp]),
    "lq_server_stats": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "lq_set_interaction_variant": (i32, [i32]),
    "lq_get_interaction_variant": (i32, []),
    "lq_grid_build_noalloc": (i32, [C.POINTER(LqParams), vp, vp, vp, vp, vp, vp]),
    "lq_engine_set_interaction_variant": (i32, [vp, i32]),
    "lq_engine_get_interaction_variant": (i32, [vp]),
    "lq_engine_neighbour_counts_stale": (i32, [vp, i32, vp]),
    "lq_engine_set_sort_interval": (i32, [vp, i32]),
    "lq_engine_set_column_order": (i32, [vp, i32]),
    "lq_engine_get_column_order": (i32, [vp]),
    "lq_engine_autotune_column_order": (i32, [vp, i32, vp, vp]),
    "lq_engine_get_sort_interval": (i32, [vp]),
    "lq_version": (C.c_char_p, []),
}

EXPORTED = sorted(_PROTOS)


def lib() -> C.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m loquat_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        msg = lib().lq_last_error()
        raise LqError(status, f"{STATUS.get(status, '?')}: {msg.decode() if msg else ''}")


def derive(par: LqParams) -> LqParams:
    p = par.copy()
    check(lib().lq_params_derive(C.byref(p)))
    return p


def r2_cut(par: LqParams) -> float:
    out = f32()
    check(lib().lq_neighbour_r2_cut(C.byref(par), C.byref(out)))
    return out.value


def ptr(a) -> int | None:
    """Address of a torch tensor or numpy array (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def host_particles(pospres, velrhop, idep, str_, stn) -> LqHostParticles:
    for a in (pospres, velrhop, idep, str_, stn):
        assert a is None or (a.dtype == np.float32 and a.flags["C_CONTIGUOUS"])
    return LqHostParticles(ptr(pospres), ptr(velrhop), ptr(idep), ptr(str_), ptr(stn))


def load_parameters(prefix: str) -> tuple[LqParams, LqVw]:
    p, v = LqParams(), LqVw()
    check(lib().lq_load_parameters(prefix.encode(), C.byref(p), C.byref(v)))
    return p, v


def save_output(directory: str, par: LqParams, vw: LqVw | None, arrs: dict, posd=None) -> None:
    """lq_save_output: arrs holds pospres/velrhop/idep/str/stn in input order."""
    keep = [np.ascontiguousarray(arrs[k], np.float32) for k in ("pospres", "velrhop", "idep", "str", "stn")]
    hp = host_particles(*keep)  # keep holds the arrays alive across the call
    pd = None if posd is None else np.ascontiguousarray(posd, np.float64)
    check(lib().lq_save_output(directory.encode(), C.byref(par), C.byref(vw) if vw is not None else None,
                               C.byref(hp), ptr(pd)))


def load_restart(prefix: str, par: LqParams, idep=None) -> np.ndarray:
    """lq_load_restart: FP64 positions (np, 3); writes the void ratio into idep[:, 2] when given."""
    posd = np.zeros((par.np, 3), np.float64)
    check(lib().lq_load_restart(prefix.encode(), C.byref(par), ptr(posd), ptr(idep)))
    return posd


def load_particles(prefix: str, par: LqParams) -> dict:
    n = par.np
    arrs = {k: np.zeros((n, w), np.float32) for k, w in
            (("pospres", 4), ("velrhop", 4), ("idep", 4), ("str", 6), ("stn", 6))}
    hp = host_particles(arrs["pospres"], arrs["velrhop"], arrs["idep"], arrs["str"], arrs["stn"])
    check(lib().lq_load_particles(prefix.encode(), C.byref(par), C.byref(hp)))
    return arrs
