"""Builds libuaamg_amd.so (HIP kernels for gfx950 + host C++ + the C ABI) in-tree.

Usage: python -m uaamg_amd.build   (or uaamg_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = PKG / "libuaamg_amd.so"

SOURCES = [
    "scene.cpp",
    "poisson_host.cpp",
    "poisson_kernels.hip",
    "pcg_kernels.hip",
    "tail_kernels.hip",
    "block_kernels.hip",
    "poisson_solver.cpp",
    "capi.cpp",
    "viscosity_host.cpp",
    "viscosity_kernels.hip",
    "viscosity_solver.cpp",
    "capi_viscosity.cpp",
    "batch_solver.cpp",
    "visc_batch.cpp",
    "server.cpp",
    "capi_server.cpp",
    "capi_packed3.cpp",
    "vdb_file.cpp",
    "capi_vdb.cpp",
    "rhs_kernels.hip",
    "rhs_device.cpp",
    "face_update_kernels.hip",
    "face_update.cpp",
    "face_weights_kernels.hip",
    "face_weights.cpp",
    "finest_kernels.hip",
    "finest_device.cpp",
    "hierarchy_kernels.hip",
    "hierarchy_device.cpp",
    "visc_volumes_kernels.hip",
    "visc_volumes.cpp",
]

ARCH = os.environ.get("UAAMG_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # the only fused multiply-adds are explicit fmaf calls (operator
        # arithmetic follows the reference term by term)
        "-ffp-contract=off",
        # no packed FP32 (v_pk_*_f32) instructions in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}",
        f"-I{CSRC}",
        "-Wall",
        "-Wno-unused-result",
    ]


def _compile(src: str) -> Path:
    s = CSRC / src
    obj = OUT_DIR / (s.stem + ".o")
    lang = ["-x", "hip"]
    cmd = [hipcc(), *lang, *common_flags(), "-c", str(s), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp), "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
