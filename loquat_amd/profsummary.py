"""Summary of rocprofv3 runs of one workload for profiles/: a kernel trace
(``--kernel-trace --stats``) and counter runs (``--pmc``: FETCH_SIZE,
WRITE_SIZE, an SQ set and a TA/TD/TCP set with GRBM_GUI_ACTIVE), each written
with ``-d DIR -o run`` (DIR/<host>/<pid>_run_results.db or DIR/run_results.db,
SQLite files with a ``kernels`` view and a ``counters_collection`` view).

HBM traffic per dispatch is FETCH_SIZE and WRITE_SIZE (KiB), each scaled by
what the counter reads for the largest device copy of the run (a copy of
--copy-mib MiB, e.g. ``bench.py --calibrate-copy 512``).  The vector L1 rate is
TCP_TOTAL_CACHE_ACCESSES / (CUs x GRBM_GUI_ACTIVE / XCDs).

usage: python -m loquat_amd.profsummary --trace D --fetch D --write D --sq D
       --ta D --match k_interaction_v3 --kernel-id ID --workload W
       --algorithmic-bytes N --copy-mib 512 --commands FILE --out PREFIX
writes PREFIX.json (read by bench.py) and PREFIX.md (tables).
"""
from __future__ import annotations

import argparse
import collections
import glob
import json
import os
import re
import sqlite3

CUS, XCDS, CLOCK_GHZ = 256, 8, 2.3


def short_name(k: str) -> str:
    k = re.sub(r"^void ", "", k)
    k = k.replace("lq::(anonymous namespace)::", "").replace("lq::", "")
    depth, out = 0, []
    for ch in k:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def db_of(d: str) -> str:
    hits = sorted(glob.glob(os.path.join(d, "**", "*results.db"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no rocprofv3 results database under {d}")
    return hits[0]


def trace_stats(d: str) -> dict:
    con = sqlite3.connect(db_of(d))
    st = collections.defaultdict(list)
    for name, s, e in con.execute("select name, start, end from kernels"):
        st[short_name(name)].append((e - s) / 1e3)
    return {k: {"calls": len(v), "total_us": sum(v), "avg_us": sum(v) / len(v)} for k, v in st.items()}


def counters(d: str) -> dict:
    """{kernel: {counter: [value per dispatch]}} (per-SE rows summed)."""
    con = sqlite3.connect(db_of(d))
    rows = con.execute("select dispatch_id, kernel_name, counter_name, sum(value) from counters_collection "
                       "group by dispatch_id, counter_name")
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for _, k, c, v in rows:
        out[short_name(k)][c].append(float(v))
    return out


def mean(v):
    return sum(v) / len(v) if v else float("nan")


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    for k in ("trace", "fetch", "write", "sq", "ta"):
        ap.add_argument("--" + k, required=True)
    ap.add_argument("--match", default="k_interaction_v3", help="kernels summarised in detail")
    ap.add_argument("--kernel-id", required=True, help="the 'kernel' key bench.py looks up")
    ap.add_argument("--workload", required=True)
    ap.add_argument("--algorithmic-bytes", type=float, required=True, help="per launch of the matched kernels")
    ap.add_argument("--copy-mib", type=float, default=512.0)
    ap.add_argument("--commands", required=True, help="text file: the commands of the runs, one per line")
    ap.add_argument("--title", default="")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()

    tr = trace_stats(a.trace)
    fe, wr, sq, ta = counters(a.fetch), counters(a.write), counters(a.sq), counters(a.ta)
    copy = "__amd_rocclr_copyBuffer"
    copy_kib = a.copy_mib * 1024.0
    ff = copy_kib / max(fe[copy]["FETCH_SIZE"]) if copy in fe else 1.0
    wf = copy_kib / max(wr[copy]["WRITE_SIZE"]) if copy in wr else 1.0
    matched = sorted(k for k in tr if a.match in k)
    kern = {}
    for k in matched:
        f_kib, w_kib = mean(fe[k]["FETCH_SIZE"]), mean(wr[k]["WRITE_SIZE"])
        tcp, grbm = mean(ta[k]["TCP_TOTAL_CACHE_ACCESSES"]), mean(ta[k]["GRBM_GUI_ACTIVE"])
        vmem = mean(sq[k]["SQ_INSTS_VMEM_RD"])
        kern[k] = {"calls": tr[k]["calls"], "avg_us": round(tr[k]["avg_us"], 2), "fetch_kib": f_kib,
                   "write_kib": w_kib, "traffic_bytes": (f_kib * ff + w_kib * wf) * 1024.0, "tcp_accesses": tcp,
                   "accesses_per_cu_per_clock": tcp / (CUS * grbm / XCDS), "lines_per_load": tcp / vmem,
                   "valu_insts": mean(sq[k]["SQ_INSTS_VALU"]), "vmem_rd_insts": vmem}
    calls = sum(v["calls"] for v in kern.values())
    wavg = lambda key: sum(v[key] * v["calls"] for v in kern.values()) / calls  # noqa: E731
    summary = {
        "kernel": a.kernel_id, "workload": a.workload,
        "commands": [ln.strip() for ln in open(a.commands) if ln.strip()],
        "fetch_factor": ff, "write_factor": wf,
        "traffic_bytes": wavg("traffic_bytes"), "avg_us": round(wavg("avg_us"), 2),
        "algorithmic_bytes": a.algorithmic_bytes,
        "launches": {k: v["calls"] for k, v in kern.items()},
        "kernels": kern,
        "l1": {"tcp_accesses_per_launch": round(wavg("tcp_accesses"), 1),
               "accesses_per_cu_per_clock": round(wavg("accesses_per_cu_per_clock"), 3),
               "lines_per_load": round(wavg("lines_per_load"), 1), "clock_ghz": CLOCK_GHZ,
               "counters": "TCP_TOTAL_CACHE_ACCESSES per dispatch / (256 CUs x GRBM_GUI_ACTIVE / 8 XCDs), "
                           "mean over the matched launches",
               "note": "the vector L1 serves about one cache-line access per CU per clock"},
    }
    with open(a.out + ".json", "w") as fh:
        json.dump(summary, fh, indent=1)

    md = [f"# {a.title or a.kernel_id}", "", f"Workload: {a.workload}.", "",
          "## Commands (from the repository root with `TMPDIR=/tmp` exported; one counter set per run)", "", "```",
          *summary["commands"], "```", "", "## Kernel trace (`--kernel-trace --stats`)", "",
          "| kernel | calls | total µs | avg µs | % |", "|---|---|---|---|---|"]
    tot = sum(v["total_us"] for v in tr.values())
    for k, v in sorted(tr.items(), key=lambda kv: -kv[1]["total_us"])[:20]:
        md.append(f"| `{k[:110]}` | {v['calls']} | {v['total_us']:.1f} | {v['avg_us']:.2f} | "
                  f"{100 * v['total_us'] / tot:.2f} |")
    md += ["", "## HBM bytes per dispatch (FETCH_SIZE, WRITE_SIZE in KiB, calibrated)", "",
           f"Calibration: the {a.copy_mib:.0f} MiB device copy (`{copy}`, the largest copy dispatch of the run) "
           f"scales FETCH_SIZE by **{ff:.3f}** and WRITE_SIZE by **{wf:.3f}**.", "",
           "| kernel | FETCH KiB | WRITE KiB | HBM read (calibrated) | HBM write (calibrated) |",
           "|---|---|---|---|---|"]
    for k in sorted(set(fe) & set(wr), key=lambda k: -mean(fe[k]["FETCH_SIZE"]))[:14]:
        f_kib, w_kib = mean(fe[k]["FETCH_SIZE"]), mean(wr[k]["WRITE_SIZE"])
        md.append(f"| `{k[:110]}` | {f_kib:.0f} | {w_kib:.0f} | {f_kib * ff * 1024 / 1e6:.1f} MB | "
                  f"{w_kib * wf * 1024 / 1e6:.1f} MB |")
    sqc = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS", "SQ_ACTIVE_INST_VALU",
           "SQ_BUSY_CU_CYCLES", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"]
    tac = ["GRBM_GUI_ACTIVE", "TA_TA_BUSY", "TA_ADDR_STALLED_BY_TC_CYCLES", "TD_TD_BUSY", "TD_TC_STALL",
           "TCP_TOTAL_CACHE_ACCESSES", "TCP_PENDING_STALL_CYCLES"]
    tac = [c for c in tac if any(c in ta[k] for k in matched)]
    md += ["", "## Instruction and memory-pipeline counters per dispatch (interaction launches)", "",
           "| kernel | " + " | ".join(sqc) + " |", "|---" * (len(sqc) + 1) + "|"]
    for k in matched:
        md.append(f"| `{k}` | " + " | ".join(f"{mean(sq[k].get(c, [])):.3g}" for c in sqc) + " |")
    md += ["", "| kernel | " + " | ".join(tac) + " |", "|---" * (len(tac) + 1) + "|"]
    for k in matched:
        md.append(f"| `{k}` | " + " | ".join(f"{mean(ta[k].get(c, [])):.3g}" for c in tac) + " |")
    md += ["", "## Reading", ""]
    for k, v in kern.items():
        cyc = mean(ta[k]["GRBM_GUI_ACTIVE"]) / XCDS
        md.append(f"* `{k}` ({v['calls']} launches): {v['avg_us']:.2f} µs per launch (trace average); calibrated "
                  f"HBM traffic {v['traffic_bytes'] / 1e6:.1f} MB per launch "
                  f"({v['traffic_bytes'] / (v['avg_us'] * 1e3):.0f} GB/s); texture address unit busy "
                  f"{100 * mean(ta[k]['TA_TA_BUSY']) / CUS / cyc:.0f} %, data unit "
                  f"{100 * mean(ta[k]['TD_TD_BUSY']) / CUS / cyc:.0f} %; VALU active in "
                  f"{100 * mean(sq[k]['SQ_ACTIVE_INST_VALU']) / mean(sq[k]['SQ_BUSY_CU_CYCLES']):.0f} % of the SQ's "
                  f"busy cycles; {v['valu_insts']:.3g} VALU and {v['vmem_rd_insts']:.3g} vector-load instructions; "
                  f"vector L1 {v['accesses_per_cu_per_clock']:.2f} accesses per CU per clock, "
                  f"{v['lines_per_load']:.1f} per load instruction.")
    md.append(f"* Per launch, mean over the {calls} matched launches: {summary['avg_us']:.1f} µs and "
              f"{summary['traffic_bytes'] / 1e6:.1f} MB of HBM traffic against {a.algorithmic_bytes / 1e6:.1f} MB "
              f"of algorithmic bytes ({summary['traffic_bytes'] / a.algorithmic_bytes:.2f}x).")
    with open(a.out + ".md", "w") as fh:
        fh.write("\n".join(md) + "\n")
    print(f"wrote {a.out}.json and {a.out}.md: {calls} launches, {summary['avg_us']:.1f} us, "
          f"{summary['traffic_bytes'] / 1e6:.1f} MB per launch")


if __name__ == "__main__":
    main()
