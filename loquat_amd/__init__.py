"""MI355X-native LOQUAT SPH time step: a C-ABI HIP library (gfx950) with a
Python front end.  `abi` binds include/loquat_amd.h; `device` holds torch
device arrays and the engine wrapper; `cases` generates projects."""
from . import abi, cases  # noqa: F401

__all__ = ["abi", "cases"]
