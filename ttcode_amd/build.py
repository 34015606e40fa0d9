"""Builds libttgpu.so (HIP, gfx950) in-tree: one object per translation unit
under ttcode_amd/csrc/, compiled in parallel, then linked.

    python -m ttcode_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libttgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # the reference evaluates without FMA contraction; keep its operation order
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
    "-I", os.path.join(HERE, "..", "include"),
]

SOURCES = ["tt_state.hip", "tt_sor.hip", "tt_phys.hip", "tt_energy.hip", "tt_hydro.hip", "tt_cnn.hip", "tt_visc.hip",
           "tt_noalloc.hip"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FDIR, "tt_driver")


def build_fortran(verbose: bool = False) -> str | None:
    """The ISO_C_BINDING layer and the Fortran step driver (needs flang)."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmd = [FLANG, "-O2", "-module-dir", mod, "-I", mod, os.path.join(FDIR, "ttgpu_mod.f90"),
           os.path.join(FDIR, "tt_driver.f90"), "-o", DRIVER, "-L", HERE, "-lttgpu",
           "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJ)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed:\n{r.stderr}")
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
