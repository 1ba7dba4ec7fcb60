"""Build libctprj.so (HIP, gfx950) in-tree with hipcc.

    python -m imgrecsrc_amd.build           # -> imgrecsrc_amd/lib/libctprj.so

Every HIP source is compiled with packed-FP32 instruction selection disabled and FP
contraction off: the forward and adjoint kernels must evaluate each footprint weight
with the same instruction sequence, so no FMA may be formed behind the code's back.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libctprj.so")
ROOT = os.path.dirname(PKG)

SOURCES = ["ctprj_kernels.hip", "ctprj_sym.hip", "bhc_kernels.hip", "ctprj_api.hip", "bhc_api.hip",
           "bhc_pnpg.hip", "bhc_sim.hip", "wavelet.hip", "tv.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    # hipcc's host pass does not know the packed-fp32 target feature; drop that noise
    noise = ("packed-fp32-ops", "not a recognized feature", "ignoring feature")
    for line in (r.stdout + r.stderr).splitlines():
        if line.strip() and not any(n in line for n in noise):
            sys.stderr.write(line + "\n")


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compiles every source (in parallel, at most `jobs` at once) and links the library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    if jobs is None:
        jobs = min(len(cmds), max(1, (os.cpu_count() or 1) // 2), 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
