"""Builds the HIP library ``libcgocean.so`` in-tree for gfx950 (MI355X).

hipcc cross-compiles without a GPU. Every HIP source is compiled with contraction off, so the
relaxation, the residual and the transfers round exactly as the reference's JVM arithmetic, and
with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libcgocean.so"

HIP_SOURCES = ["runtime.hip", "unet.hip", "coarse_net.hip"]
CXX_SOURCES = ["host_builders.cpp", "save_sections.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    """Compiles every source (the translation units side by side) and links ``libcgocean.so``;
    always rebuilds."""
    from concurrent.futures import ThreadPoolExecutor

    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    jobs = []
    for src in CXX_SOURCES:
        jobs.append(["g++", *COMMON, "-pthread", "-c", str(CSRC / src), "-o", str(obj_dir / (src + ".o"))])
    for src in HIP_SOURCES:
        jobs.append([hipcc(), *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj_dir / (src + ".o"))])
    objs = [obj_dir / (src + ".o") for src in CXX_SOURCES + HIP_SOURCES]
    with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
        list(pool.map(run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    run([hipcc(), "-shared", "-fPIC", "-pthread", "--offload-arch=gfx950", *map(str, objs), "-lz", "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
