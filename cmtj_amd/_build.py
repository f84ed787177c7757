"""Builds libcmtj_amd.so (HIP, gfx950) in-tree.

Each kernel instantiation unit compiles to its own object in parallel, then
hipcc links the shared library next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libcmtj_amd.so"
UNITS = ["engine.hip", "k_f64.hip", "k_f64e.hip", "k_f32.hip", "k_spec.hip", "k_ll_f64.hip", "k_ll_f32.hip",
         "k_ll_spec.hip", "k_exact.hip", "k_cl.hip", "k_cl64.hip", "k_stack.hip", "k_dp.hip", "k_group.hip"]
ARCH = os.environ.get("CMTJ_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off (hipcc would otherwise form them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return exe


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    hipcc = _hipcc()

    def compile_unit(unit: str) -> Path:
        src = CSRC / unit
        obj = objdir / (unit + ".o")
        cmd = [hipcc, *FLAGS, "-c", str(src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed for {unit}")
        if verbose:
            sys.stderr.write(r.stderr)
        return obj

    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_unit, UNITS))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
