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
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baseline's OpenMP threads stay on neighbouring cores (read by the
# OpenMP runtime when the oracle library loads).
os.environ.setdefault("OMP_PROC_BIND", "close")

METRIC = "M particle-steps/s at 1M particles (1 GPU); ensemble p50/p99 vs plain loop"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3       # published FP32 vector peak (packed FP32 instructions)
FP32_VALU_PEAK_TFLOPS = 78.6   # non-packed FP32 vector (packed FP32 is disabled in the build)
FLOPS_PER_PAIR = 80            # algorithmic FP32 flops per neighbour pair
BYTES_READ_PER_PARTICLE = 64   # x (FP64, 24) + v (12) + rho (4) + stress (24)
BYTES_WRITE_PER_MATERIAL = 52  # a (12) + drho (4) + spin (12) + strain rate (24)
# With the integration fused into the interaction (variant 19 on runs without
# artificial pressure or a dynamic time step) no rates are written; instead,
# per material particle and half step, beyond what the scan reads:
# predictor: reads ep, void ratio (8), strain (24), p (4), table position
# (12); writes x (FP64, 24), FP32 position + p (16), v + rho (16), stress (24)
# = 128 B; corrector: reads the step's start x (24), v + rho (16), stress
# (24), strain (24), ep, void ratio (8), p (4), table position (12); writes
# x (24), position + p (16), v + rho (16), ep, void ratio (8), stress (24),
# strain (24) = 224 B.  Mean per launch: 176 B.
BYTES_INTEG_PER_MATERIAL = 176
FUSED_VARIANTS = (19,)
BYTES_PER_CELL = 8             # cell_beg + cell_end
WORKLOAD_ID = "column_3d dx=5mm npm=1024000"  # the key of this workload in profiles/*.json
# Template arguments of k_interaction_v3 per variant (3 and up): pairs per
# iteration, masked gathers, list capacity, candidates per round, waves per
# workgroup, partial-flush size, column order, lanes per particle, staging.
V3_SHAPES = {3: (1, False, 40, 4, 4, 0, 0, 1, False), 4: (2, False, 40, 4, 4, 0, 0, 1, False),
             5: (1, True, 40, 4, 4, 0, 0, 1, False), 6: (2, True, 40, 4, 4, 0, 0, 1, False),
             7: (2, True, 32, 4, 4, 0, 0, 1, False), 8: (2, True, 48, 4, 4, 0, 0, 1, False),
             9: (2, True, 40, 8, 4, 0, 0, 1, False), 10: (2, True, 40, 4, 2, 0, 0, 1, False),
             11: (2, True, 40, 4, 1, 0, 0, 1, False), 12: (2, True, 40, 4, 1, 12, 1, 1, False),
             13: (2, True, 40, 4, 1, 16, 2, 1, False), 14: (2, True, 40, 4, 1, 8, 2, 1, False),
             15: (2, True, 40, 4, 1, 0, 2, 1, False), 16: (2, True, 40, 4, 1, 16, 2, 2, False),
             17: (2, True, 32, 4, 1, 16, 2, 1, True), 18: (2, False, 40, 4, 1, 16, 2, 1, False),
             19: (2, False, 32, 4, 1, 16, 2, 1, True)}


def v3_name(shape: tuple, stale: bool, fuse: int = 0) -> str:
    """Demangled name of one k_interaction_v3 instance (fuse 1: with the
    predictor fused in, 2: the corrector)."""
    b, masked, cap, u, wpb, pf, order, lpp, stg = shape
    f = lambda x: "true" if x else "false"  # noqa: E731
    return (f"k_interaction_v3<DescTable, {b}, {f(masked)}, false, {f(stale)}, {cap}, {u}, {wpb}, {pf}, {order}, "
            f"{lpp}, {f(stg)}, {fuse}>")


KERNEL_NAMES = {0: "k_interaction_direct", 1: "k_interaction_lds", 2: "k_interaction_unrolled",
                **{v: v3_name(sh, False) for v, sh in V3_SHAPES.items()}}


# Between a run's sorts both half steps scan its kept cell table: variants 11
# and up in their own shape, the others in variant 6's shape (variant 10's
# workgroup size for 10); the predictor half of a sort step scans the fresh
# table with the variant.
def stale_kernel(variant: int) -> str:
    if variant >= 11:
        return v3_name(V3_SHAPES[variant], True)
    return v3_name((2, True, 40, 4, 2 if variant == 10 else 4, 0, 0, 1, False), True)


def roofline_kernel(variant: int) -> str:
    if variant in FUSED_VARIANTS:
        sh = V3_SHAPES[variant]
        return (f"{v3_name(sh, True, 1)} + {v3_name(sh, True, 2)} (kept tables, predictor and corrector "
                f"fused in) + {v3_name(sh, False, 1)} (fresh table at sorts)")
    return f"{stale_kernel(variant)} (kept tables) + {KERNEL_NAMES[variant]} (fresh table at sorts)"


def bytes_per_launch(variant: int, np_: int, npm: int, ncell: int) -> int:
    """Algorithmic HBM bytes of one interaction launch (with the fused
    integration where the variant has it)."""
    per_mat = BYTES_INTEG_PER_MATERIAL if variant in FUSED_VARIANTS else BYTES_WRITE_PER_MATERIAL
    return np_ * BYTES_READ_PER_PARTICLE + npm * per_mat + ncell * BYTES_PER_CELL


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_info() -> dict:
    """CPU of this host, as nproc and lscpu report it (read before the GPU is
    touched)."""
    info = {"nproc": len(os.sched_getaffinity(0)), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                             "CPU max MHz"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(case, steps: int, host: dict) -> dict:
    """FP64 oracle (C++/OpenMP) on this host: initial sort, 2 warm-up steps,
    then `steps` timed steps of the same workload on all threads, and one
    more step on one thread."""
    from oracle import lq_oracle as O

    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    O.run(p, case.vw, st, 2, 0)
    t0 = time.perf_counter()
    O.run(p, case.vw, st, steps, 2)
    dt = time.perf_counter() - t0
    threads = O.num_threads()
    O.set_num_threads(1)
    t1 = time.perf_counter()
    O.run(p, case.vw, st, 1, 2 + steps)
    dt1 = time.perf_counter() - t1
    O.set_num_threads(threads)
    return {
        "value": case.npm * steps / dt / 1e6,
        "unit": "M particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle, same 1.55 M-particle column: 2 warm-up + {steps} timed full steps "
                  f"({dt:.1f} s), OpenMP threads={threads} bound close, {host.get('Model name', '')}",
        "s_per_step": dt / steps,
        "one_thread_s_per_step": dt1,
        "host": host,
    }


def cpu_baseline_c1(steps: int, one_thread_steps: int) -> dict:
    """The plane-strain sand column (5,000 material particles), all of its
    steps on the FP64 oracle with every OpenMP thread, and a bounded sample on
    one thread."""
    from loquat_amd import cases
    from oracle import lq_oracle as O

    case = cases.column_2d_small(nsteps=steps)
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    t0 = time.perf_counter()
    O.run(p, case.vw, st, steps, 0)
    dt = time.perf_counter() - t0
    threads = O.num_threads()
    O.set_num_threads(1)
    st1 = O.OState.from_case(case)
    O.initial_sort(p, st1)
    t1 = time.perf_counter()
    O.run(p, case.vw, st1, one_thread_steps, 0)
    dt1 = time.perf_counter() - t1
    O.set_num_threads(threads)
    return {"workload": f"plane-strain sand column 50x100x1 at dx=2 mm, npm={case.npm}, np={case.np}, FP64",
            "steps": steps, "seconds": round(dt, 3), "threads": threads,
            "M_particle_steps_s": round(case.npm * steps / dt / 1e6, 4),
            "one_thread": {"steps": one_thread_steps, "s_per_step": round(dt1 / one_thread_steps, 6),
                           "M_particle_steps_s": round(case.npm * one_thread_steps / dt1 / 1e6, 4)}}


def profile_entry(kernel: str) -> tuple[dict | None, str | None]:
    """The committed rocprofv3 summary (profiles/*.json) of `kernel` on this
    workload, and its path, or (None, None)."""
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    if not os.path.isdir(d):
        return None, None
    for f in sorted(os.listdir(d)):
        if not f.endswith(".json"):
            continue
        with open(os.path.join(d, f)) as fh:
            j = json.load(fh)
        if j.get("kernel") == kernel and j.get("workload") == WORKLOAD_ID:
            return j, "profiles/" + f
    return None, None


def measured_traffic(kernel: str) -> tuple[float | None, str | None]:
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 summary of the same command (FETCH_SIZE + WRITE_SIZE per
    dispatch, each scaled by what the counter reads for a device copy of
    known size), or None when no summary covers this kernel."""
    j, src = profile_entry(kernel)
    return (float(j["traffic_bytes"]), src) if j else (None, None)


def measured_launch_us(kernel: str) -> float | None:
    """Mean launch time of `kernel` on this workload in the committed
    rocprofv3 kernel trace (profiles/*.json), beside the live event timing."""
    j, _ = profile_entry(kernel)
    return float(j["avg_us"]) if j and "avg_us" in j else None


def measured_l1(kernel: str) -> dict | None:
    """Vector L1 (TCP) cache-line accesses per launch and per CU per clock of
    `kernel` on this workload, from the committed counter summary."""
    j, src = profile_entry(kernel)
    if not j or "l1" not in j:
        return None
    return dict(j["l1"], source=src)


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


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, float), q)) if len(xs) else None


def ensemble_section(nruns: int, steps: int, warm: int, seed: int = 0, plain_loop: bool = True) -> dict:
    """256 independent 3D columns (npm log-uniform 2k..60k, aspect U[0.5,4],
    phi U[20,40] deg, c U[0,5] kPa): all runs in one engine (one launch per
    stage, graph-replayed) against the plain loop (the same engine code, one
    run at a time; skipped without plain_loop).  Bounded sample: `steps` timed
    steps per run."""
    import torch

    from loquat_amd import cases
    from loquat_amd.device import Engine

    runs = cases.ensemble(nruns=nruns, seed=seed, nsteps=steps + warm)
    npm_tot = sum(c.npm for c in runs)
    # batched
    eng = Engine(sum(c.np for c in runs), sum(c.derived().ncell for c in runs), len(runs))
    ids = [eng.add(c) for c in runs]
    eng.step(warm)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.sync()
    t_batch = time.perf_counter() - t0
    # timed work actually done (a failed run stops early)
    done = [eng.info(r)["steps"] for r in ids]
    work = sum(c.npm * max(0, k - warm) for c, k in zip(runs, done))
    failed = sum(k < steps + warm for k in done)
    eng.close()
    torch.cuda.empty_cache()
    if not plain_loop:
        return {"workload": f"{nruns} independent 3D column collapses (sum npm {npm_tot})",
                "steps": steps, "warmup": warm, "batched_M_particle_steps_s": round(work / t_batch / 1e6, 3),
                "ms_per_step": round(1e3 * t_batch / steps, 4), "failed_runs": failed}
    # plain loop: one run at a time, each a batch of one
    big = max(runs, key=lambda c: c.np)
    eng = Engine(big.np, max(c.derived().ncell for c in runs), 1)
    per_run, completion, clock, work_plain = [], [], 0.0, 0
    for c in runs:
        rid = eng.add(c)
        eng.step(warm)
        eng.sync()
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        dt = time.perf_counter() - t0
        work_plain += c.npm * max(0, eng.info(rid)["steps"] - warm)
        eng.remove(rid)
        per_run.append(dt)
        clock += dt
        completion.append(clock)
    eng.close()
    batched = work / t_batch / 1e6
    plain = work_plain / sum(per_run) / 1e6
    return {
        "workload": f"{nruns} independent 3D column collapses, npm log-uniform 2k..60k (sum {npm_tot}), "
                    f"DP phi U[20,40] deg, c U[0,5] kPa, Shepard every 20 steps",
        "sample": f"{steps} timed steps per run (+{warm} warm-up) of the 5,000",
        "npm_total": npm_tot,
        "batched_M_particle_steps_s": round(batched, 3),
        "plain_loop_M_particle_steps_s": round(plain, 3),
        "speedup_vs_plain_loop": round(batched / plain, 3),
        "failed_runs": failed,
        "batched_completion_s": {"p50": round(t_batch, 4), "p99": round(t_batch, 4),
                                 "note": "every run has the same step count, so all runs of the batch "
                                         "complete together: p50 = p99 is the measured completion time"},
        "plain_loop_completion_s": {"p50": round(_pct(completion, 50), 4), "p99": round(_pct(completion, 99), 4)},
    }


def served_section(nreq: int, step_scale: float, seed: int = 0) -> dict:
    """Poisson arrivals (20 runs/s) of column (Drucker-Prager) and embankment
    (von Wolffersdorff) runs, 50/50, material particles log-normal in
    [5k, 500k], 2,000-8,000 steps scaled by step_scale; admitted by HBM
    footprint into one engine and batched every step (C++ request loop).
    The plain loop serves the same arrivals one at a time, FIFO."""
    import math

    import numpy as np
    import torch

    from loquat_amd import cases
    from loquat_amd.device import Engine
    from loquat_amd.serve import Server

    reqs = cases.served_mix(nreq, step_scale, seed)
    np_cap = sum(c.np for _, c, _ in reqs)
    nc_cap = sum(c.derived().ncell for _, c, _ in reqs)
    srv = Server(np_cap, nc_cap, 256, 288 * 10**9, steps_per_tick=25)
    done_at, ids, fin = {}, {}, {}
    t0 = time.perf_counter()
    nxt = 0
    busy = 0.0   # time inside ticks with runs resident: the server's device-busy time
    while len(done_at) < nreq:
        now = time.perf_counter() - t0
        while nxt < nreq and reqs[nxt][0] <= now:
            ids[srv.submit(reqs[nxt][1], reqs[nxt][2])] = nxt
            nxt += 1
        if srv.stats()["running"] == 0 and srv.stats()["queued"] == 0:
            if nxt < nreq:
                time.sleep(max(0.0, min(0.01, reqs[nxt][0] - now)))
            continue
        tt = time.perf_counter()
        finished = srv.tick()
        busy += time.perf_counter() - tt
        for rid in finished:
            done_at[ids[rid]] = time.perf_counter() - t0
            fin[ids[rid]] = srv.info(rid)
            log(f"served: {len(done_at)}/{nreq} done at {done_at[ids[rid]]:.1f} s")
    makespan = time.perf_counter() - t0
    srv.close()
    torch.cuda.empty_cache()
    lat = [done_at[i] - reqs[i][0] for i in range(nreq)]
    # work actually done: a run that fails stops early and is credited only
    # with the steps it took
    work = sum(reqs[i][1].npm * fin[i]["steps"] for i in range(nreq))
    failed = sum(fin[i]["state"] != "done" for i in range(nreq))
    # plain loop: FIFO, one run at a time on the same engine code; each run's
    # time covers the same work a tick does for it (upload and first sort,
    # its steps, removal), so both throughputs are work / device-busy time
    big = max(reqs, key=lambda r: r[1].np)[1]
    eng = Engine(big.np, max(c.derived().ncell for _, c, _ in reqs), 1)
    clock, lat_plain, busy_plain, work_plain = 0.0, [], 0.0, 0
    by_size = {"over_200k": [0, 0.0], "under_200k": [0, 0.0]}  # plain loop: work, time per size class
    for a, c, s in reqs:
        t1 = time.perf_counter()
        rid = eng.add(c, s)
        eng.step(s)
        eng.sync()
        steps_done = eng.info(rid)["steps"]
        eng.remove(rid)
        dt = time.perf_counter() - t1
        work_plain += c.npm * steps_done
        busy_plain += dt
        cls = by_size["over_200k" if c.npm >= 2e5 else "under_200k"]
        cls[0] += c.npm * steps_done
        cls[1] += dt
        clock = max(clock, a) + dt
        lat_plain.append(clock - a)
        log(f"plain loop: {len(lat_plain)}/{nreq} ({c.npm} material particles, {s} steps) in {dt:.2f} s")
    eng.close()
    served = work / busy / 1e6
    plain = work_plain / busy_plain / 1e6
    big_work = sum(c.npm * fin[i]["steps"] for i, (_, c, _) in enumerate(reqs) if c.npm >= 2e5)
    return {
        "workload": f"{nreq} requests, Poisson 20/s, column DP / slope von Wolffersdorff 50/50, "
                    f"npm log-normal clipped [5k,500k], steps U[2000,8000] x {step_scale}",
        "latency_s": {"p50": round(_pct(lat, 50), 4), "p99": round(_pct(lat, 99), 4)},
        "plain_loop_latency_s": {"p50": round(_pct(lat_plain, 50), 4), "p99": round(_pct(lat_plain, 99), 4)},
        "busy_M_particle_steps_s": round(served, 3),
        "plain_loop_M_particle_steps_s": round(plain, 3),
        "speedup_vs_plain_loop": round(served / plain, 3),
        "throughput_definition": "material particle-steps / device-busy time for both (server: time inside "
                                 "ticks with runs resident; plain loop: add + steps + remove of each run)",
        "makespan_M_particle_steps_s": round(work / makespan / 1e6, 3),
        "work_share_runs_over_200k": round(big_work / max(work, 1), 4),
        "plain_loop_M_particle_steps_s_by_size": {k: round(w / max(t, 1e-9) / 1e6, 3) for k, (w, t) in by_size.items()},
        "np_per_npm": round(sum(c.np for _, c, _ in reqs) / sum(c.npm for _, c, _ in reqs), 3),
        "failed_runs": failed,
        "makespan_s": round(makespan, 3),
        "busy_s": round(busy, 3),
    }


def slope_section(steps: int, npm_target: float = 4.0e6, autotune: bool = True, cell_order: int = 4) -> dict:
    """30-degree embankment, ~4 M material particles, von Wolffersdorff with a
    seeded +-5% void-ratio field.  With autotune, the in-cell order (starting
    from cell_order) and then the interaction variant are tuned on the run's
    state before the 5 untimed steps."""
    import torch

    from loquat_amd import cases
    from loquat_amd.device import Engine

    c = cases.slope_vw(npm_target=npm_target, nsteps=steps + 5)
    eng = Engine(c.np, c.derived().ncell, 1)
    eng.set_cell_order(cell_order)
    rid = eng.add(c)
    tuned = None
    if autotune:
        bits, _ = eng.autotune_cell_order(3)
        rows, _ = eng.autotune_column_order(3)
        var, _ = eng.autotune(3)
        tuned = {"cell_order_zbits": bits, "column_strip_rows": rows, "variant": var}
    eng.step(5)
    eng.sync()
    s = torch.cuda.ExternalStream(eng.stream())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    eng.step(steps)
    e1.record(s)
    eng.sync()
    ms = e0.elapsed_time(e1)
    done = eng.info(rid)["steps"] - 5
    eng.close()
    if done != steps:
        raise SystemExit(f"bench: the embankment ran {done} of {steps} timed steps (the run failed)")
    torch.cuda.empty_cache()
    return {"workload": f"30-degree embankment 1.0 x 3.0 x 0.5 m, npm={c.npm}, np={c.np}, "
                        f"dx={c.meta['dx'] * 1e3:.2f} mm, von Wolffersdorff",
            "M_particle_steps_s": round(c.npm * steps / (ms * 1e-3) / 1e6, 3),
            "ms_per_step": round(ms / steps, 4), "steps": steps, "autotune": tuned}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=None, help="interaction kernel variant (without --autotune)")
    ap.add_argument("--autotune", type=int, default=1, help="1: time the in-cell orders (from --cell-order, "
                    "kept unless another is 1 %% faster) and every interaction variant on the workload's state "
                    "and run the fastest")
    ap.add_argument("--ensemble-runs", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-step-scale", type=float, default=0.05)
    ap.add_argument("--slope-steps", type=int, default=20)
    ap.add_argument("--no-extra", action="store_true", help="skip the ensemble / served / slope sections")
    ap.add_argument("--slope-only", action="store_true",
                    help="time only the embankment section and print its JSON (for profiling that workload)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the batched ensemble (no plain loop): --ensemble-steps timed after --warmup")
    ap.add_argument("--served-only", action="store_true",
                    help="run only the served section (with --serve-requests / --serve-step-scale) and print its JSON")
    ap.add_argument("--sort-interval", type=int, default=None, metavar="K",
                    help="re-sort each run every K of its steps (default: the engine's)")
    ap.add_argument("--column-order", type=int, default=None, metavar="T",
                    help="(x, y) column order of the engine's sort: strips of T x-rows (default: the engine's)")
    ap.add_argument("--cell-order", type=int, default=4, metavar="B",
                    help="in-cell order of the step's sort: 0 = by cell only, B > 0 = 2^B z-bins per cell")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="after the run, 5 device copies of MB MiB (reference kernel for HBM byte counters)")
    ap.add_argument("--c1-steps", type=int, default=20000,
                    help="steps of the plane-strain column timed on the FP64 oracle (all of them: 20000)")
    ap.add_argument("--c1-one-thread-steps", type=int, default=1000)
    ap.add_argument("--no-ref-order", action="store_true",
                    help="skip timing the column again in the reference's in-cell order")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the SPH step of one run lives on one device "
              "(ensembles shard by whole runs with no communication); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    host = host_info()   # before anything touches the GPU

    import torch

    from loquat_amd import abi, cases
    from loquat_amd.device import Engine

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if a.variant is None:
        a.variant = abi.DEFAULT_INTERACTION_VARIANT
    # the calling thread's variant: engines created from here on start with it
    abi.check(abi.lib().lq_set_interaction_variant(a.variant))
    if a.slope_only:
        sec = slope_section(a.slope_steps, autotune=bool(a.autotune), cell_order=a.cell_order)
        print(json.dumps({"slope_vw": sec}), flush=True)
        return
    if a.ensemble_only:
        print(json.dumps({"ensemble": ensemble_section(a.ensemble_runs, a.ensemble_steps, a.warmup,
                                                       plain_loop=False)}), flush=True)
        return
    if a.served_only:
        print(json.dumps({"served": served_section(a.serve_requests, a.serve_step_scale)}), flush=True)
        return

    # room for the timed steps twice: in the tuned order and in the reference order
    case = cases.column_3d(nsteps=2 * (a.warmup + a.steps + a.profile_steps) + 1)
    p = case.derived()
    log(f"workload: np={case.np} npm={case.npm} npb={p.npb} ncell={p.ncell} dt={p.dt:.4e}")

    pairs = pair_count(case)
    log(f"neighbour pairs of the initial state: {pairs} ({pairs / case.npm:.2f} per material particle)")

    eng = Engine(case.np, p.ncell, 1)
    eng.set_cell_order(a.cell_order)
    if a.column_order is not None:
        eng.set_column_order(a.column_order)
    if a.sort_interval is not None:
        eng.set_sort_interval(a.sort_interval)
    column_order = eng.column_order()
    sort_interval = eng.sort_interval()
    rid = eng.add(case)
    tuned = None
    cell_order = a.cell_order
    if a.autotune:
        cell_order, oms = eng.autotune_cell_order(3)
        column_order, cms = eng.autotune_column_order(3)
        best, ms = eng.autotune(3)
        tuned = {"variant": best, "kernel": KERNEL_NAMES[best], "launch_ms": [round(m, 4) for m in ms],
                 "cell_order_zbits": cell_order,
                 "ms_per_cell_order": [round(m, 4) for m in oms],
                 "column_strip_rows": column_order,
                 "ms_per_column_order": dict(zip((0, 2, 4, 8, 16), (round(m, 4) for m in cms)))}
        log(f"autotune: cell order {cell_order} (sort + Adami + interaction ms per z-bin bits 0..: "
            f"{tuned['ms_per_cell_order']}), column strips {column_order} ({tuned['ms_per_column_order']}), "
            f"variant {best}, launch ms per variant {tuned['launch_ms']}")
    variant = tuned["variant"] if tuned else a.variant
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
    if eng.info(rid)["steps"] != a.warmup + a.steps:
        raise SystemExit(f"bench: the column ran {eng.info(rid)['steps']} of {a.warmup + a.steps} steps "
                         "(the run failed); no throughput to report")

    # Per-stage device time (HIP events around each stage, direct launches,
    # same stream and state, right after the timed loop).
    stage = eng.profile(a.profile_steps)
    info = eng.info(rid)
    # The same run timed again in the reference's sort order (by cell id
    # only, SortParticleByCellIndex: no z-bins, x-major columns): the cost of
    # that exact order.
    ref_order = None
    if not a.no_ref_order:
        eng.set_cell_order(0)
        eng.set_column_order(0)
        eng.step(a.warmup)
        eng.sync()
        ev2, ev3 = hip.Event(enable_timing=True), hip.Event(enable_timing=True)
        ev2.record(s)
        eng.step(a.steps)
        ev3.record(s)
        eng.sync()
        ms_ref = ev2.elapsed_time(ev3) / a.steps
        stage_ref = eng.profile(a.profile_steps)
        ref_order = {"value": round(case.npm / (ms_ref * 1e-3) / 1e6, 3), "ms_per_step": round(ms_ref, 4),
                     "cell_order_zbits": 0, "column_strip_rows": 0,
                     "stage_ms_per_step": {k: round(v, 4) for k, v in
                                           zip(("grid", "adami", "interaction", "integrate", "shepard"), stage_ref)}}
        log(f"reference in-cell order: {ref_order}")
        if eng.info(rid)["steps"] < info["steps"] + a.warmup + a.steps:
            raise SystemExit("bench: the column failed in the reference in-cell order")
    eng.close()
    inter_launch_ms = stage[2] / 2.0   # two interaction launches per step
    log("stage ms/step: grid %.3f adami %.3f interaction %.3f integrate %.3f shepard %.3f" % tuple(stage))

    bytes_launch = bytes_per_launch(variant, case.np, case.npm, p.ncell)
    flops_launch = pairs * FLOPS_PER_PAIR
    gbs = bytes_launch / (inter_launch_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(roofline_kernel(variant))
    tflops = flops_launch / (inter_launch_ms * 1e-3) / 1e12
    # the binding bound: the larger of the launch's HBM time and its FP32 time
    # at the published peak (VALU for these algorithmic counts)
    t_hbm = bytes_launch / (HBM_PEAK_GBS * 1e9)
    t_valu = flops_launch / (FP32_PEAK_TFLOPS * 1e12)
    t_valu_np = flops_launch / (FP32_VALU_PEAK_TFLOPS * 1e12)
    launch_s = inter_launch_ms * 1e-3
    trace_us = measured_launch_us(roofline_kernel(variant))

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
            "cell_order_zbits": cell_order,
            "column_strip_rows": column_order,
            "sort_interval": sort_interval,
            "np_steps_per_s_M": round(case.np * a.steps / (gpu_ms * 1e-3) / 1e6, 3),
            "wall_ms_per_step": round(wall * 1e3 / a.steps, 4),
            "steps_done": info["steps"],
        },
        "roofline": {
            "kernel": roofline_kernel(variant),
            "measured_by": "HIP events around the interaction launches on the engine stream (lq_engine_profile); "
                           "launch_ms is the mean of the step's two launches",
            # the bound whose time at the published peak is the longer
            "bound": "hbm" if t_hbm >= t_valu else "valu",
            "achieved": round(gbs, 2) if t_hbm >= t_valu else round(tflops, 3),
            "peak": HBM_PEAK_GBS if t_hbm >= t_valu else FP32_PEAK_TFLOPS,
            "unit": "GB/s" if t_hbm >= t_valu else "TFLOP/s",
            "frac": round(max(t_hbm, t_valu) / launch_s, 5),
            "frac_definition": "max(bytes / 8 TB/s, flops / 157.3 TFLOP/s) / launch time (with the integration "
                               "fused in, its bytes are counted, its flops are not)",
            "valu": {"achieved_tflops": round(tflops, 3), "peak_tflops": FP32_PEAK_TFLOPS,
                     "frac": round(t_valu / launch_s, 5)},
            "fused_integration": variant in FUSED_VARIANTS,
            "frac_nonpacked": round(max(t_hbm, t_valu_np) / launch_s, 5),
            "peak_nonpacked": FP32_VALU_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "launch_ms": round(inter_launch_ms, 5),
            "launch_us_rocprof_trace": trace_us,
            "flops_per_launch": flops_launch,
            "pairs": pairs,
            "bytes_per_launch": bytes_launch,
            "hbm": {"achieved_gbs": round(gbs, 2), "peak_gbs": HBM_PEAK_GBS,
                    "frac": round(gbs / HBM_PEAK_GBS, 5)},
            "l1": measured_l1(roofline_kernel(variant)),
        },
        "stage_ms_per_step": {k: round(v, 4) for k, v in
                              zip(("grid", "adami", "interaction", "integrate", "shepard"), stage)},
        "reference_in_cell_order": ref_order,
        "autotune": tuned,
        "cpu_baseline": None,
    }
    torch.cuda.empty_cache()
    if a.calibrate_copy > 0:
        n = a.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = n * 4
        del src, dst
        torch.cuda.empty_cache()
    if not a.no_extra:
        t = time.perf_counter()
        out["ensemble"] = ensemble_section(a.ensemble_runs, a.ensemble_steps, 10)
        log(f"ensemble: {out['ensemble']} ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        out["served"] = served_section(a.serve_requests, a.serve_step_scale)
        log(f"served: {out['served']} ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        out["slope_vw"] = slope_section(a.slope_steps, autotune=bool(a.autotune), cell_order=a.cell_order)
        log(f"slope: {out['slope_vw']} ({time.perf_counter() - t:.1f} s)")
    if not a.no_cpu:
        log(f"CPU baseline: FP64 oracle, {a.cpu_steps} steps ...")
        out["cpu_baseline"] = cpu_baseline(cases.column_3d(nsteps=a.cpu_steps + 3), a.cpu_steps, host)
        if a.c1_steps > 0:
            t = time.perf_counter()
            out["cpu_baseline_c1"] = cpu_baseline_c1(a.c1_steps, a.c1_one_thread_steps)
            log(f"CPU baseline C1: {out['cpu_baseline_c1']} ({time.perf_counter() - t:.1f} s)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
