"""Writes c3_slope_trajectory.json: the von Wolffersdorff embankment of the
slope workload at reduced resolution (0.2 M material particles requested,
dx = 1.75 cm, 323,843 particles with walls), 3,000 steps (0.105 s),
computed by the FP64 oracle from the generator's seeded initial state: runout, centre of mass, height and
free-surface profile (as for the columns) plus the material displacement
field's mean and 99th percentile, which a slope that barely moves makes the
sharper check.  About 15 minutes on 8 cores.

    python tests/golden/make_c3_trajectory.py [--steps 3000] [--chunk 250]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from loquat_amd import cases  # noqa: E402
from oracle import lq_oracle as O  # noqa: E402
from metrics import collapse_metrics, displacement_metrics  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--chunk", type=int, default=250)
    a = ap.parse_args()
    case = cases.slope_vw(npm_target=2e5, nsteps=a.steps)
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    t0 = time.time()
    done = 0
    while done < a.steps:
        k = min(a.chunk, a.steps - done)
        O.run(p, case.vw, st, k, done)
        done += k
        print(f"{done}/{a.steps} steps, {time.time() - t0:.0f} s", flush=True)
    posd = O.unpermute(st).posd
    m = collapse_metrics(case, posd)
    m.update(displacement_metrics(case, posd))
    m.update(case="slope_vw", npm_target=2e5, steps=a.steps, oracle_seconds=round(time.time() - t0, 1),
             threads=O.num_threads())
    with open(os.path.join(HERE, "c3_slope_trajectory.json"), "w") as f:
        json.dump(m, f, indent=1)
    print(json.dumps({k: v for k, v in m.items() if k != "profile"}))


if __name__ == "__main__":
    main()
