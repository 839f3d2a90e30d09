"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_err(gpu, ref, floor: float = 0.0) -> float:
    """max_i |gpu - ref| / max(max_i |ref|, floor), per field."""
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    if ref.size == 0:
        return 0.0
    scale = max(float(np.abs(ref).max()), floor)
    if scale == 0.0:
        return float(np.abs(gpu - ref).max())
    return float(np.abs(gpu - ref).max() / scale)
