"""Benchmark of the SPH time step on one MI355X.

Workload: 3D granular column collapse, 80 x 80 x 160 = 1,024,000 material
particles at dx = 5 mm (529,254 wall particles, 949,105 cells), Drucker-Prager,
Shepard filter every 20 steps, FP32 sums with FP64 positions.  One step is the
reference's full step: two half steps (grid, Adami, interaction, integration)
plus the Shepard pass every 20th step.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-steps S] [--no-cpu]

Prints one JSON line: value = material particle-steps per second (millions),
timed over K steps replayed from hipGraphs with the state resident in HBM;
roofline of the interaction kernel from HIP events on the engine stream; the
FP64 oracle timed on this host as the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M particle-steps/s at 1M particles (1 GPU); ensemble p50/p99 vs plain loop"
HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 78.6   # non-packed FP32 vector (packed FP32 is disabled in the build)
FLOPS_PER_PAIR = 80            # algorithmic FP32 flops per neighbour pair
BYTES_READ_PER_PARTICLE = 64   # x (FP64, 24) + v (12) + rho (4) + stress (24)
BYTES_WRITE_PER_MATERIAL = 52  # a (12) + drho (4) + spin (12) + strain rate (24)
BYTES_PER_CELL = 8             # cell_beg + cell_end


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(case, steps: int) -> dict:
    """FP64 oracle (C++/OpenMP) on this host: initial sort, 1 warm-up step,
    then `steps` timed steps of the same workload."""
    from oracle import lq_oracle as O

    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    O.run(p, case.vw, st, 1, 0)
    t0 = time.perf_counter()
    O.run(p, case.vw, st, steps, 1)
    dt = time.perf_counter() - t0
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": case.npm * steps / dt / 1e6,
        "unit": "M particle-steps/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"FP64 oracle, same 1.55 M-particle column: 1 warm-up + {steps} timed full steps "
                  f"({dt:.1f} s), OpenMP threads={O.num_threads()}, {cpu_model}",
        "s_per_step": dt / steps,
    }


def pair_count(case) -> int:
    """Neighbour pairs (i material, j != i, q < 2) of the workload's initial
    state, counted by the HIP interaction through the module ABI."""
    import torch

    from loquat_amd import device as dev

    p = case.derived()
    src = dev.DeviceState.from_case(case)
    dst = dev.DeviceState(case.np)
    g = dev.Grid(case.np, p.ncell)
    g.build(p, src)
    g.permute(case.np, src, dst)
    beg, end = g.table_ptrs()
    rt = dev.DeviceRates(case.np)
    nbr = torch.zeros(case.np, dtype=torch.int32, device="cuda")
    dev.interaction(p, dst, rt, beg, end, nbr)
    torch.cuda.synchronize()
    n = int(nbr[p.npb:].to(torch.int64).sum().item())
    g.close()
    del src, dst, rt, nbr
    torch.cuda.empty_cache()
    return n


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=1, help="interaction kernel variant")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the SPH step of one run lives on one device "
              "(ensembles shard by whole runs with no communication); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from loquat_amd import abi, cases
    from loquat_amd.device import Engine

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    abi.check(abi.lib().lq_set_interaction_variant(a.variant))

    case = cases.column_3d(nsteps=a.warmup + a.steps + a.profile_steps + 1)
    p = case.derived()
    log(f"workload: np={case.np} npm={case.npm} npb={p.npb} ncell={p.ncell} dt={p.dt:.4e}")

    pairs = pair_count(case)
    log(f"neighbour pairs of the initial state: {pairs} ({pairs / case.npm:.2f} per material particle)")

    eng = Engine(case.np, p.ncell, 1)
    rid = eng.add(case)
    eng.step(a.warmup)
    eng.sync()
    torch.cuda.synchronize()
    log(f"warm-up {a.warmup} steps done")

    hip = torch.cuda
    s = torch.cuda.ExternalStream(eng.stream())
    ev0, ev1 = hip.Event(enable_timing=True), hip.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(s)
    eng.step(a.steps)
    ev1.record(s)
    eng.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / a.steps
    log(f"timed {a.steps} steps: {gpu_ms:.2f} ms on the engine stream, {wall * 1e3:.2f} ms wall")

    # Per-stage device time (HIP events around each stage, direct launches,
    # same stream and state, right after the timed loop).
    stage = eng.profile(a.profile_steps)
    info = eng.info(rid)
    eng.close()
    inter_launch_ms = stage[2] / 2.0   # two interaction launches per step
    log("stage ms/step: grid %.3f adami %.3f interaction %.3f integrate %.3f shepard %.3f" % tuple(stage))

    bytes_launch = case.np * BYTES_READ_PER_PARTICLE + case.npm * BYTES_WRITE_PER_MATERIAL + p.ncell * BYTES_PER_CELL
    flops_launch = pairs * FLOPS_PER_PAIR
    gbs = bytes_launch / (inter_launch_ms * 1e-3) / 1e9
    tflops = flops_launch / (inter_launch_ms * 1e-3) / 1e12

    value = case.npm * a.steps / (gpu_ms * 1e-3) / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "3D granular column collapse, 80x80x160 material particles at dx=5 mm "
                        "(0.4x0.4x0.8 m column, 2.0x0.4 m box, 3 Adami wall layers), Drucker-Prager "
                        "phi=30 deg, Shepard every 20 steps, FP32 sums + FP64 positions",
            "npm": case.npm, "np": case.np, "ncell": p.ncell, "dx": 0.005,
            "particle_step": "one material particle advanced one full predictor-corrector step",
            "np_steps_per_s_M": round(case.np * a.steps / (gpu_ms * 1e-3) / 1e6, 3),
            "wall_ms_per_step": round(wall * 1e3 / a.steps, 4),
            "steps_done": info["steps"],
        },
        "roofline": {
            "kernel": "k_interaction_lds" if a.variant == 1 else "k_interaction_direct",
            "bound": "hbm",
            "achieved": round(gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 5),
            "traffic": None,
            "launch_ms": round(inter_launch_ms, 5),
            "bytes_per_launch": bytes_launch,
            "valu": {
                "flops_per_launch": flops_launch,
                "pairs": pairs,
                "achieved_tflops": round(tflops, 3),
                "peak_tflops": FP32_VALU_PEAK_TFLOPS,
                "frac": round(tflops / FP32_VALU_PEAK_TFLOPS, 5),
            },
        },
        "stage_ms_per_step": {k: round(v, 4) for k, v in
                              zip(("grid", "adami", "interaction", "integrate", "shepard"), stage)},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        log(f"CPU baseline: FP64 oracle, {a.cpu_steps} steps ...")
        out["cpu_baseline"] = cpu_baseline(cases.column_3d(nsteps=a.cpu_steps + 1), a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
