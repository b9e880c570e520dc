"""Build lsmps_vpm_v2_amd/libvpmfmm.so for gfx950 with hipcc (in-tree).

Usage: ``python -m lsmps_vpm_v2_amd.build`` (or ``python lsmps_vpm_v2_amd/build.py``).
Every source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled; tree.hip additionally disables FMA contraction so cell
geometry rounds exactly as the reference's FP64 arithmetic; so do advect.hip
and fmm3d.hip (whose sums opt back in to contraction block by block).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libvpmfmm.so")
OBJ = os.path.join(HERE, "_build")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-x", "hip",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{os.path.join(HERE, '..', 'include')}",
]

SOURCES = {
    "prim.hip": [],
    "tree.hip": ["-ffp-contract=off"],
    "fmm64.hip": [],
    "fmm32.hip": [],
    "select.hip": [],
    "advect.hip": ["-ffp-contract=off"],
    "fmm3d.hip": ["-ffp-contract=off"],
    "api.cpp": [],
}


def _compile(name: str, extra: list[str]) -> str:
    src = os.path.join(CSRC, name)
    obj = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    names = [n for n in SOURCES if os.path.exists(os.path.join(CSRC, n))]
    with ThreadPoolExecutor(max_workers=min(8, len(names))) as ex:
        objs = list(ex.map(lambda n: _compile(n, SOURCES[n]), names))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
