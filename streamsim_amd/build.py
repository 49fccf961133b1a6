"""Build the HIP engine library ``streamsim_amd/libfacet_hip.so`` for gfx950.

Every translation unit under ``csrc/`` is compiled with hipcc in parallel into
``streamsim_amd/_build/`` and linked in-tree.  The ray-tracing unit is built
with FP contraction off (the bit-exact ray replay depends on every float
expression rounding as written); all units keep packed-FP32 instructions off.

Usage: ``python -m streamsim_amd.build [--jobs N]``
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libfacet_hip.so"
CLI = PKG / "facet_engine"
ARCH = os.environ.get("FH_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]
PER_FILE = {
    "raytrace.hip": ["-ffp-contract=off"],
    "raster.hip": ["-ffp-contract=off"],
    "optics.hip": ["-ffp-contract=off"],
}
SOURCES = ["abi.hip", "raytrace.hip", "spmm.hip", "eb.hip", "mlp.hip", "raster.hip", "optics.hip", "bvh.cpp"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return exe


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [hipcc(), "-c", *COMMON, *PER_FILE.get(src.name, []), str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    # command-line engine (host C++ over the C-ABI), rpath to this folder
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(PKG / "host" / "facet_engine.cpp"),
           f"-L{PKG}", "-lfacet_hip", "-lz", "-Wl,-rpath,$ORIGIN", "-o", str(CLI)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"facet_engine build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {LIB} and {CLI}", file=sys.stderr)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args()
    build(args.jobs)


if __name__ == "__main__":
    main()
