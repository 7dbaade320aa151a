"""Build libmcrt_hip.so (gfx950) in-tree with hipcc.

    python montecarloradiativetransfer.jl_amd/build.py [-v]

Every source under csrc/ is compiled as HIP for gfx950 with packed-FP32
instructions disabled, then linked into ``libmcrt_hip.so`` next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libmcrt_hip.so"
# (source, object name, extra flags): mcrt_coop.hip is compiled once per part
# (its MCRT_COOP_PART sections) so that its transport kernels build in parallel.
SOURCES = [(src, src.rsplit(".", 1)[0], []) for src in (
    "mcrt_tables.cpp", "mcrt_api.cpp", "mcrt_batch.cpp", "mcrt_sched.cpp", "mcrt_serve.cpp",
    "mcrt_plain.hip", "mcrt_static.hip")] + [
    ("mcrt_coop.hip", f"mcrt_coop_{part}", [f"-DMCRT_COOP_PART={part}"]) for part in range(1, 6)]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-x", "hip",
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... ({cmd[-1]})")
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)


def build_library(verbose: bool = False, jobs: int = 0) -> Path:
    """Compile every source and link the shared library; returns its path."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found: cannot build libmcrt_hip.so")
    BUILD.mkdir(exist_ok=True)
    if jobs <= 0:
        jobs = max(1, min(8, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                          else (os.cpu_count() or 4)))
    objs = []
    cmds = []
    for src, name, flags in SOURCES:
        obj = BUILD / (name + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, *flags, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
