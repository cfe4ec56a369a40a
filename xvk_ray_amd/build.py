"""Build step of the native libraries (no GPU needed: hipcc cross-compiles gfx950).

Produces, in-tree:
  xvk_ray_amd/lib/libxr.so       the HIP layer behind include/xr_api.h
  xvk_ray_amd/lib/libxr_host.so  the C++ host frame driver behind include/xr_host.h

Usage: python -m xvk_ray_amd.build   (or xvk_ray_amd.build.build())
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
HOST = PKG / "host"
LIB = PKG / "lib"
OBJ = PKG / "_build"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -ffp-contract=off + correctly rounded FP32 div/sqrt: every expression in the
# kernels evaluates to the same bits as on the host (explicit fmaf() where an
# FMA is wanted).  Packed-FP32 instructions are disabled for gfx950.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{ROOT / 'include'}", f"-I{CSRC}"]

HIP_SOURCES = ["xr_api.hip", "xr_kernels.hip", "xr_bvh_build.hip", "xr_mlp.hip"]
HOST_SOURCES = ["scene.cpp", "gltf.cpp", "png.cpp", "png_decode.cpp", "camera.cpp", "driver.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile_hip(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(out)])
    return out


def _compile_host(src: str) -> Path:
    out = OBJ / ("host_" + Path(src).stem + ".o")
    _run(["g++", *CXX_FLAGS, "-c", str(HOST / src), "-o", str(out)])
    return out


def build(verbose: bool = True) -> None:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        hip_objs = list(ex.map(_compile_hip, HIP_SOURCES))
        host_objs = list(ex.map(_compile_host, [s for s in HOST_SOURCES if (HOST / s).exists()]))
    libxr = LIB / "libxr.so"
    tmp = libxr.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, hip_objs)])
    os.replace(tmp, libxr)
    if host_objs:
        libh = LIB / "libxr_host.so"
        tmp = libh.with_suffix(".so.tmp")
        _run(["g++", "-shared", "-fPIC", "-o", str(tmp), *map(str, host_objs),
              f"-L{LIB}", "-lxr", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, libh)
    if verbose:
        print(f"built {libxr}" + (f" and {LIB / 'libxr_host.so'}" if host_objs else ""))


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
