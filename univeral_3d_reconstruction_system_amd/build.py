"""Builds libu3d_hip.so (HIP kernels + C ABI) for gfx950, in-tree.

Every translation unit is compiled with floating-point contraction off: the projection, tile
keys, compositing, loss and optimiser arithmetic must match the reference trainer's operation
order bit for bit. Packed FP32 instructions are disabled for all device code.

Usage: python -m univeral_3d_reconstruction_system_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libu3d_hip.so")
ARCH = os.environ.get("U3D_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels_project.hip", "kernels_sort.hip", "kernels_raster.hip", "kernels_loss.hip",
           "kernels_adam.hip", "kernels_points.hip", "kernels_mlp.hip", "kernels_view.hip", "api.cpp",
           "batch.cpp", "initnet.cpp", "ply.cpp", "view.cpp"]

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-x", "hip", *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    # -z defs: a symbol left undefined (objects from mismatched sources) fails the link here
    # instead of the library's first load
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host(jobs)
    if verbose:
        print(f"built {LIB} and {SPLAT_LIB}")
    return LIB


HOST_DIR = os.path.join(HERE, "host")
SPLAT_LIB = os.path.join(OUT_DIR, "libu3d_splat.so")
HOST_SOURCES = ["Camera.cpp", "ThreadPool.cpp", "GaussianModel.cpp", "device.cpp", "Rasterizer.cpp",
                "Loss.cpp", "Adam.cpp", "Trainer.cpp", "SplatPly.cpp"]


def _compile_host(src: str) -> str:
    obj = os.path.join(OBJ_DIR, "host_" + os.path.splitext(src)[0] + ".o")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra",
           "-I", os.path.join(os.path.dirname(HERE), "include"), "-I", os.path.join(HOST_DIR, "include"),
           "-c", os.path.join(HOST_DIR, "src", src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_host(jobs: int | None = None) -> str:
    """libu3d_splat.so: the reference's C++ class API (u3d::Rasterizer, Trainer, ...) as thin
    shims over libu3d_hip's C ABI; headers in host/include."""
    jobs = jobs or min(len(HOST_SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_host, HOST_SOURCES))
    tmp = SPLAT_LIB + ".tmp"
    cmd = ["g++", "-shared", "-fPIC", "-o", tmp, *objs, "-L", OUT_DIR, "-lu3d_hip",
           "-Wl,-rpath,$ORIGIN", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SPLAT_LIB)
    return SPLAT_LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(a.jobs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
