"""Writes c1_slab_trajectory.json: final runout, centre of mass, height and
free-surface profile of the plane-strain sand column as an 8-particle slab
between side walls (50 x 8 x 100 material particles at dx = 2 mm, 20,000
steps), computed by the FP64 oracle from the generator's seeded initial
state.  --perturb EPS additionally runs the oracle from positions jittered
by +-EPS and records the metric spread (the collapse is chaotic).

    python tests/golden/make_c1_trajectory.py [--steps 20000] [--perturb 1e-9]
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
from metrics import collapse_metrics  # noqa: E402


def run(case, eps: float, seed: int):
    p = case.derived()
    st = O.OState.from_case(case)
    if eps:
        st.x[p.npb:] += np.random.default_rng(seed).uniform(-eps, eps, p.npm)
    O.initial_sort(p, st)
    t = time.time()
    O.run(p, case.vw, st, case.nsteps)
    return collapse_metrics(case, O.unpermute(st).posd), time.time() - t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--perturb", type=float, default=0.0)
    a = ap.parse_args()
    case = cases.column_2d_slab(nsteps=a.steps)
    m, el = run(case, 0.0, 0)
    m.update(case="column_2d_slab", steps=a.steps, seed=0, oracle_seconds=round(el, 1), threads=O.num_threads())
    if a.perturb:
        m2, _ = run(case, a.perturb, 1)
        m["perturbed"] = {"eps_m": a.perturb, **{k: m2[k] for k in ("runout", "com_x", "com_z", "height")},
                          "profile": m2["profile"]}
    with open(os.path.join(HERE, "c1_slab_trajectory.json"), "w") as f:
        json.dump(m, f, indent=1)
    print(json.dumps({k: v for k, v in m.items() if k not in ("profile", "perturbed")}))


if __name__ == "__main__":
    main()
