"""Order-independent full-run metrics of a column collapse."""
import numpy as np


def collapse_metrics(case, posd) -> dict:
    """runout: max x of material particles within 2 dx of the floor, measured
    from the wall; profile: max height above the floor of the material in
    x-bins of width 2 dx from the wall (NaN for empty bins)."""
    dx = case.meta["dx"]
    x0, z0 = case.meta["wall_x"], case.meta["floor_z"]
    npb = case.par.npb
    x = posd[npb:, 0] - x0
    z = posd[npb:, 2] - z0
    ok = np.isfinite(x) & np.isfinite(z)
    x, z = x[ok], z[ok]
    base = z < 2 * dx
    runout = float(x[base].max())
    nb = int(np.ceil((x.max() + 1e-12) / (2 * dx)))
    b = np.clip((x / (2 * dx)).astype(int), 0, nb - 1)
    prof = np.full(nb, np.nan)
    for i in range(nb):
        sel = z[b == i]
        if sel.size:
            prof[i] = sel.max()
    return {"runout": runout, "height": float(z.max()), "bin": 2 * dx,
            "profile": [None if np.isnan(v) else float(v) for v in prof]}
