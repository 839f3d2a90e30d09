"""Device idle time between the kernels of engine steps, from a rocprofv3
kernel trace (``rocprofv3 --kernel-trace -d DIR -o run -- python3 ...``
writes DIR/<host>/<pid>_run_results.db or DIR/run_results.db, an SQLite file
with a ``kernels`` view of name, start and end in ns).

Steps are delimited by the engine's k_step_begin launches; for every complete
step after the first `skip` (at most `steps` of them), the span (one
k_step_begin to the next) is compared with the summed durations of the
kernels launched in it.  Idle time is what a persistent kernel or a fused
step could remove at most.

usage: python -m loquat_amd.tracegaps RUN_RESULTS.db [--skip N] [--steps M]
"""
from __future__ import annotations

import argparse
import sqlite3
import statistics


def step_gaps(rows: list[tuple[str, int, int]], skip: int = 0, nsteps: int | None = None) -> dict:
    """rows: (kernel name, start ns, end ns), any order.  Returns the mean
    span, busy and idle time per step (us), the kernels per step and the
    largest single gap (us) over the complete steps after `skip` (the first
    `nsteps` of them when given)."""
    rows = sorted(rows, key=lambda r: r[1])
    begins = [i for i, r in enumerate(rows) if "k_step_begin" in r[0]]
    steps = []
    pairs = list(zip(begins, begins[1:]))[skip:]
    for a, b in pairs[:nsteps] if nsteps is not None else pairs:
        ks = rows[a:b]
        span = rows[b][1] - rows[a][1]
        busy = sum(e - s for _, s, e in ks)
        gaps = [max(0, ks[i + 1][1] - ks[i][2]) for i in range(len(ks) - 1)] + [max(0, rows[b][1] - ks[-1][2])]
        steps.append((span, busy, len(ks), max(gaps)))
    if not steps:
        raise ValueError("fewer than two k_step_begin launches after the skipped steps")
    span = statistics.mean(s[0] for s in steps) / 1e3
    busy = statistics.mean(s[1] for s in steps) / 1e3
    return {"steps": len(steps), "span_us": span, "busy_us": busy, "idle_us": span - busy,
            "busy_frac": busy / span, "kernels_per_step": statistics.mean(s[2] for s in steps),
            "largest_gap_us": max(s[3] for s in steps) / 1e3}


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("db")
    ap.add_argument("--skip", type=int, default=0, help="steps to leave out at the start (warm-up)")
    ap.add_argument("--steps", type=int, default=None, help="steps to summarise after those")
    a = ap.parse_args()
    con = sqlite3.connect(a.db)
    rows = list(con.execute("select name, start, end from kernels"))
    r = step_gaps(rows, a.skip, a.steps)
    print(f"{r['steps']} steps: span {r['span_us']:.1f} us/step, kernels busy {r['busy_us']:.1f} us/step "
          f"({100 * r['busy_frac']:.2f} %), idle {r['idle_us']:.1f} us/step over {r['kernels_per_step']:.1f} "
          f"kernels; largest gap {r['largest_gap_us']:.1f} us")


if __name__ == "__main__":
    main()
