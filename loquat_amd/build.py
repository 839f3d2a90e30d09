"""Builds the HIP library (gfx950) and the FP64 oracle in-tree.

    python -m loquat_amd.build          # both
    python -m loquat_amd.build --lib    # HIP library only

The library lands in loquat_amd/lib/libloquat_amd.so, objects in
loquat_amd/_build/; the oracle (test infrastructure) in oracle/liblq_oracle.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libloquat_amd.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liblq_oracle.so")

ARCH = os.environ.get("LOQUAT_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["lq_grid.hip", "lq_physics.hip", "lq_engine.hip"]
CXX_SOURCES = ["lq_io.cpp", "lq_server.cpp"]

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    # hipcc's host pass does not know the device target feature: harmless.
    noise = "-packed-fp32-ops"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l and l.strip())
    if verbose and err:
        sys.stderr.write(err + "\n")


def build_lib(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(OBJ, s.replace(".cpp", ".o"))
        objs.append(o)
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
                     "-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = True) -> str:
    _run(["make", "-s", "-B", "-C", ORACLE_DIR, "liblq_oracle.so"], verbose)
    return ORACLE_LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="HIP library only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    ap.add_argument("-q", "--quiet", action="store_true")
    a = ap.parse_args()
    v = not a.quiet
    if not a.oracle:
        print(build_lib(v))
    if not a.lib:
        print(build_oracle(v))


if __name__ == "__main__":
    main()
