"""Build the gfx950 library (and, for the tests, the CPU checkers) in-tree.

    python -m vebisph_amd.build            # HIP library + Fortran interface + oracle (+ reference)

Outputs (git-ignored, shipped with the tree):
    vebisph_amd/libvebisph_hip.so     the product: HIP kernels + the C ABI of include/vebisph_gpu.h
    vebisph_amd/libvebisph_f.so       Fortran module vebisph_gpu (iso_c_binding) + div_free_gpu
    oracle/liboracle.so, oracle/_ref/libvbref.so   test checkers (oracle/Makefile)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vebisph_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libvebisph_hip.so")
FLIB = os.path.join(PKG, "libvebisph_f.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # every a*b+c stays two rounded operations, as in the reference's FP64 code
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_hip(jobs: int = 8, out: str = LIB, defines: tuple = ()) -> str:
    """Compile csrc/*.hip into `out`.  `defines` (e.g. "VB_SPMV_BLOCK=256") override the launch
    constants of vb_common.h, for launch-shape sweeps into libraries outside the package."""
    bdir = BUILD if out == LIB else out + ".build"
    os.makedirs(bdir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    dflags = [f"-D{d}" for d in defines]

    def one(src):
        obj = os.path.join(bdir, src[:-4] + ".o")
        _run([HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_fortran() -> str | None:
    """iso_c_binding interface module and the Fortran div_free_gpu driver (flang)."""
    if not os.path.exists(FLANG):
        return None
    src = os.path.join(PKG, "fortran", "vebisph_gpu.F90")
    if not os.path.exists(src):
        return None
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, "vebisph_gpu_f.o")
    _run([FLANG, "-cpp", "-O2", "-fPIC", "-module-dir", BUILD, "-c", src, "-o", obj])
    tmp = FLIB + ".tmp"
    _run([FLANG, "-shared", "-fPIC", "-o", tmp, obj, f"-L{PKG}", "-lvebisph_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    return FLIB


REFIO_LIB = os.path.join(PKG, "libvebisph_refio.so")


def build_refio() -> str:
    """libvebisph_refio.so: readers/writers of the reference's run files (host C++, no GPU)."""
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-fPIC", "-shared",
          f"-I{os.path.join(ROOT, 'include')}", os.path.join(PKG, "cpp", "vb_refio.cpp"),
          "-o", REFIO_LIB + ".tmp"])
    os.replace(REFIO_LIB + ".tmp", REFIO_LIB)
    return REFIO_LIB


def build_cpp_driver() -> str:
    """vb_drive: the stand-alone C++ driver of the C ABI (vebisph_amd/cpp/vb_drive.cpp)."""
    out = os.path.join(PKG, "vb_drive")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(PKG, "cpp", "vb_drive.cpp"), "-o", out + ".tmp", f"-L{PKG}", "-lvebisph_hip",
          "-lvebisph_refio", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_checkers() -> None:
    """oracle/liboracle.so always; oracle/_ref (the reference's step and its case generator) when the
    reference sources are present."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if os.path.isdir("/root/reference/source"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "refnewt", "gen"])


def build_all() -> None:
    build_hip()
    build_fortran()
    build_refio()
    build_cpp_driver()
    build_checkers()


if __name__ == "__main__":
    build_all()
    print("built", LIB, file=sys.stderr)
