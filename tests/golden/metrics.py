"""Order-independent full-run metrics of a column collapse.

A granular collapse is chaotic: perturbing the initial positions by 1e-9 m
(far below FP32 resolution) moves the single farthest particle by several
per cent.  These metrics use percentiles and means instead of extremes, so
that two correct implementations can be compared at the 1 % level."""
import numpy as np


def collapse_metrics(case, posd) -> dict:
    """runout: 99th percentile of x (from the wall) of material particles
    within 2 dx of the floor; com_x/com_z: centre of mass from wall/floor;
    height: 99th percentile of the height above the floor; profile: 90th
    percentile height in x-bins of width 10 dx from the wall (None for bins
    holding fewer than 20 particles)."""
    dx = case.meta["dx"]
    x0, z0 = case.meta["wall_x"], case.meta["floor_z"]
    npb = case.par.npb
    x = posd[npb:, 0] - x0
    z = posd[npb:, 2] - z0
    ok = np.isfinite(x) & np.isfinite(z)
    x, z = x[ok], z[ok]
    base = z < 2 * dx
    w = 10 * dx
    nb = int(np.ceil(x.max() / w)) + 1
    b = np.clip((x / w).astype(int), 0, nb - 1)
    prof = []
    for i in range(nb):
        sel = z[b == i]
        prof.append(float(np.percentile(sel, 90)) if sel.size >= 20 else None)
    return {"runout": float(np.percentile(x[base], 99)), "com_x": float(x.mean()), "com_z": float(z.mean()),
            "height": float(np.percentile(z, 99)), "bin": w, "profile": prof,
            "below_floor": int(np.sum(z < 0)), "count": int(x.size)}


def displacement_metrics(case, posd) -> dict:
    """Mean and 99th percentile of the material particles' displacement from
    the case's initial positions (m), over the particles still finite."""
    npb = case.par.npb
    d = np.linalg.norm(np.asarray(posd[npb:], np.float64) - case.posd[npb:], axis=1)
    d = d[np.isfinite(d)]
    return {"disp_mean": float(d.mean()), "disp_p99": float(np.percentile(d, 99))}
