"""lq_grid_build_noalloc (the grid build for callers that cannot allocate)
against lq_grid_build on a grid handle: cell keys, sort permutation and cell
table bit-identical, on two calls in a row, the second with a smaller input
(and a different grid), both with particles outside the grid."""
import ctypes as C

import numpy as np
import pytest

from loquat_amd import abi, cases
from oracle import lq_oracle as O


def _shuffled_with_outside(case, seed):
    """The case in a random input order, with a few material particles
    outside the grid (they hash to ncell - 1)."""
    rng = np.random.default_rng(seed)
    p = case.derived()
    perm = np.concatenate([np.arange(p.npb), p.npb + rng.permutation(case.npm)])
    for k in ("pospres", "velrhop", "idep", "str", "stn", "posd"):
        setattr(case, k, np.ascontiguousarray(getattr(case, k)[perm]))
    out = p.npb + rng.choice(case.npm, 5, replace=False)
    case.posd[out[:3], 0] = -0.5
    case.posd[out[3:], 2] = p.domain_size_z + 1.0
    case.pospres[:, :3] = case.posd.astype(np.float32)
    return case


def _noalloc(p, ds):
    import torch
    from loquat_amd.device import stream_ptr
    out = {k: torch.full((n,), 7, dtype=torch.int32, device="cuda") for k, n in
           (("part_cel", ds.n), ("part_idx", ds.n), ("cell_beg", p.ncell), ("cell_end", p.ncell))}
    abi.check(abi.lib().lq_grid_build_noalloc(C.byref(p), abi.ptr(ds.pospres), abi.ptr(out["part_cel"]),
                                              abi.ptr(out["part_idx"]), abi.ptr(out["cell_beg"]),
                                              abi.ptr(out["cell_end"]), stream_ptr()))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _with_handle(p, ds):
    from loquat_amd.device import Grid
    g = Grid(ds.n, p.ncell)
    g.build(p, ds)
    got = g.arrays(ds.n, p.ncell)
    g.close()
    return got


@pytest.mark.gpu
def test_grid_build_noalloc_equals_grid_build(gpu):
    import torch
    from loquat_amd.device import DeviceState
    big = _shuffled_with_outside(cases.perturb(cases.small_column_3d(nx=60, ny=30, nz=90, seed=3), seed=4), 5)
    small = _shuffled_with_outside(cases.perturb(cases.small_column_3d(seed=6), seed=7), 8)
    assert big.np > 4 * small.np and big.derived().ncell != small.derived().ncell
    for case in (big, small):
        p = case.derived()
        ds = DeviceState.from_case(case)
        print(f"\nlq_grid_build_noalloc: np={case.np} ncell={p.ncell}", flush=True)
        got = _noalloc(p, ds)
        ref = _with_handle(p, ds)
        for k in ("part_cel", "part_idx", "cell_beg", "cell_end"):
            assert np.array_equal(got[k], ref[k]), k
        # and the oracle's grid (stable sort of the FP32 cell hash)
        ora = O.grid_build(p, O.OState.from_case(case))
        for k in ("part_cel", "part_idx", "cell_beg", "cell_end"):
            assert np.array_equal(got[k], ora[k]), k
        assert (got["part_cel"] == p.ncell - 1).sum() == 5
        del ds
        torch.cuda.empty_cache()


@pytest.mark.gpu
def test_grid_build_noalloc_rejects_oversize(gpu):
    import torch
    from loquat_amd.device import stream_ptr
    p = cases.small_column_3d(seed=3).derived()
    n = abi.NOALLOC_MAX_PARTICLES + 1
    p.np, p.npb, p.npm = n, 0, n
    buf = torch.zeros(4 * n, dtype=torch.float32, device="cuda")
    idx = torch.zeros(n, dtype=torch.int32, device="cuda")
    tab = torch.zeros(p.ncell, dtype=torch.int32, device="cuda")
    st = abi.lib().lq_grid_build_noalloc(C.byref(p), abi.ptr(buf), abi.ptr(idx), abi.ptr(idx), abi.ptr(tab),
                                         abi.ptr(tab), stream_ptr())
    assert st == 8 and b"LQ_NOALLOC_MAX_PARTICLES" in abi.lib().lq_last_error()
