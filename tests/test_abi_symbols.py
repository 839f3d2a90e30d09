"""The C-ABI library loads and exports every entry point the header declares
(no compute call: runs without a GPU)."""
import ctypes as C
import os
import re

from loquat_amd import abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "loquat_amd.h")


def declared() -> set[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(lq_[a-z0-9_]+)\s*\(", text))


def test_header_declares_entry_points():
    names = declared()
    assert {"lq_grid_build", "lq_particle_interaction", "lq_adami_boundary", "lq_predictor", "lq_corrector",
            "lq_field_smoothing", "lq_engine_step"} <= names


def test_library_exports_every_declared_symbol():
    lib = abi.lib()
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert declared() == set(abi.EXPORTED)


def test_struct_sizes():
    assert C.sizeof(abi.LqParams) == 208   # include/types.h:89-155
    assert C.sizeof(abi.LqVw) == 32


def test_version_and_error_without_gpu():
    lib = abi.lib()
    assert lib.lq_version().startswith(b"loquat_amd")
    # argument validation runs on the host and reports through lq_last_error
    assert lib.lq_engine_create(0, 0, 0, None) == 1
    assert len(lib.lq_last_error()) > 0
    assert lib.lq_set_interaction_variant(abi.INTERACTION_VARIANTS) == 1
    assert lib.lq_get_interaction_variant() in range(abi.INTERACTION_VARIANTS)


def test_particle_capacity_limit_rejected_before_allocation():
    """Gathers use 32-bit byte offsets (24-byte elements at most): capacities
    above LQ_MAX_PARTICLES are refused with LQ_ERR_CAPACITY before any device
    allocation, so this runs without a GPU."""
    import ctypes as C
    from loquat_amd import abi
    lib = abi.lib()
    assert abi.MAX_PARTICLES * 24 < 2 ** 32 <= (abi.MAX_PARTICLES + 1) * 24
    h = C.c_void_p()
    assert lib.lq_engine_create(abi.MAX_PARTICLES + 1, 1000, 1, C.byref(h)) == 8
    assert lib.lq_grid_create(abi.MAX_PARTICLES + 1, 1000, C.byref(h)) == 8
    par = abi.LqParams()
    par.np = abi.MAX_PARTICLES + 1
    par.npm = par.np
    # module calls check the particle count before anything else
    assert lib.lq_particle_interaction(C.byref(par), None, None, None, None, None, None) == 8
