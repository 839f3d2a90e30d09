"""Writes c1_trajectory.json: final runout and free-surface profile of the
quasi-2D sand column (50 x 100 x 1 particles at dx = 2 mm, 20,000 steps)
computed by the FP64 oracle from the generator's seeded initial state.

    python tests/golden/make_c1_trajectory.py [--steps 20000]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from loquat_amd import cases  # noqa: E402
from oracle import lq_oracle as O  # noqa: E402
from metrics import collapse_metrics  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    a = ap.parse_args()
    case = cases.column_2d_small(nsteps=a.steps)
    p = case.derived()
    st = O.OState.from_case(case)
    O.initial_sort(p, st)
    t = time.time()
    O.run(p, case.vw, st, a.steps)
    el = time.time() - t
    u = O.unpermute(st)
    m = collapse_metrics(case, u.posd)
    m.update(steps=a.steps, seed=0, oracle_seconds=round(el, 1), threads=O.num_threads())
    with open(os.path.join(HERE, "c1_trajectory.json"), "w") as f:
        json.dump(m, f, indent=1)
    print(json.dumps({k: v for k, v in m.items() if k != "profile"}))


if __name__ == "__main__":
    main()
