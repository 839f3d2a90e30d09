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


def fma32(a, b, c) -> np.ndarray:
    """Correctly rounded FP32 fused multiply-add a * b + c, elementwise.  The
    FP32 product is exact in FP64; the FP64 sum is corrected by its exact
    error term (two-sum) where it falls exactly halfway between two FP32
    values, the one case in which rounding the FP64 sum to FP32 could differ
    from rounding the exact sum."""
    a64 = np.asarray(a, np.float32).astype(np.float64)
    b64 = np.asarray(b, np.float32).astype(np.float64)
    c64 = np.asarray(c, np.float32).astype(np.float64)
    p = a64 * b64
    s = p + c64
    bb = s - p
    err = (p - (s - bb)) + (c64 - bb)
    r = s.astype(np.float32)
    up = np.nextafter(r, np.float32(np.inf))
    dn = np.nextafter(r, np.float32(-np.inf))
    at_up = s == (r.astype(np.float64) + up.astype(np.float64)) * 0.5
    at_dn = s == (r.astype(np.float64) + dn.astype(np.float64)) * 0.5
    r = np.where(at_up & (err > 0), up, r)
    r = np.where(at_dn & (err < 0), dn, r)
    return r


def r2_fp32(d: np.ndarray) -> np.ndarray:
    """The neighbour predicate's squared distance of FP32 differences d (..., 3):
    fma(dz, dz, fma(dy, dy, dx * dx)) with dx * dx rounded to FP32."""
    d = np.asarray(d, np.float32)
    xx = d[..., 0] * d[..., 0]
    return fma32(d[..., 2], d[..., 2], fma32(d[..., 1], d[..., 1], xx))
