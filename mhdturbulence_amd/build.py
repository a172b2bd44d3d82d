"""Build the HIP library (libmhd_amd.so) and the C++ drop-in host library in-tree.

    python -m mhdturbulence_amd.build          # compile for gfx950
    python -m mhdturbulence_amd.build -D MHD_XY_DPP=0 --out /tmp/lib_nodpp.so   # a variant;
        # load it with MHD_AMD_LIB=/tmp/lib_nodpp.so

Each .hip source is compiled to an object with hipcc for gfx950, then linked into
mhdturbulence_amd/libmhd_amd.so.  The drop-in host layer (mhdturbulence_amd/host/)
is compiled with g++ against include/mhd_amd.h into libmhd_kh.so plus the example
driver mhd_kh_main, so that an unmodified src_cpp_kokkos_host-style step loop links
against the reference's function names.

    python -m mhdturbulence_amd.build --calib   # only the measurement tool

builds calib/mhd_calib.hip into the executable mhd_calib (HBM stream rates and FP64 issue
rate of the GPU it runs on).  It is a measurement tool, not part of the library, and
build() does not compile it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmhd_amd.so")
KH_LIB = os.path.join(PKG, "libmhd_kh.so")
KH_MAIN = os.path.join(PKG, "mhd_kh_main")
KH_MAIN_256 = os.path.join(PKG, "mhd_kh_main_256")
CALIB = os.path.join(PKG, "mhd_calib")

ARCH = os.environ.get("MHD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 arithmetic stays off: hipcc emits only scalar FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

SOURCES = ["mhd_kernels.hip", "mhd_fused.hip", "mhd_capi.hip", "mhd_batch.hip"]
# Compilation units: (source, object name, extra flags).  mhd_fused.hip is built twice: its
# host side and FP32 kernels, and its FP64 kernels with the instruction-level-parallelism
# scheduler (the FP64 march 1 % faster at 256^3; the FP32 kernels ran 4-5 % slower with it).
UNITS = [
    ("mhd_kernels.hip", "mhd_kernels", []),
    ("mhd_fused.hip", "mhd_fused", ["-DMHD_FUSED_UNIT=1"]),
    ("mhd_fused.hip", "mhd_fused64", ["-DMHD_FUSED_UNIT=2", "-mllvm", "-amdgpu-sched-strategy=max-ilp"]),
    ("mhd_capi.hip", "mhd_capi", []),
    ("mhd_batch.hip", "mhd_batch", []),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(unit, defines=(), build_dir: str = BUILD) -> str:
    src, name, extra = unit
    obj = os.path.join(build_dir, name + ".o")
    _run([HIPCC, *HIP_FLAGS, *extra, *[f"-D{d}" for d in defines], "-I", os.path.join(ROOT, "include"), "-c",
          os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip(verbose: bool = True, defines=(), out: str = LIB) -> str:
    """Compile the HIP sources (optionally with -D defines, e.g. tuning switches) into `out`."""
    build_dir = BUILD if out == LIB else out + ".objs"
    os.makedirs(build_dir, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u[0]))]
    with cf.ThreadPoolExecutor(max_workers=min(5, len(units))) as ex:
        objs = list(ex.map(lambda u: _compile(u, defines, build_dir), units))
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_host(verbose: bool = True) -> str:
    """The src_cpp_kokkos_host-compatible C++ layer over the C ABI."""
    if not os.path.isdir(HOST):
        return ""
    cxx = os.environ.get("CXX", "g++")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", HOST]
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]
    srcs = [os.path.join(HOST, f) for f in ("mhd.cpp", "boundary.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return ""
    tmp = KH_LIB + ".tmp"
    _run([cxx, *flags, *inc, "-shared", "-o", tmp, *srcs, f"-L{PKG}", "-lmhd_amd", f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, KH_LIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        _run([cxx, *flags, *inc, "-o", KH_MAIN, main, f"-L{PKG}", "-lmhd_kh", "-lmhd_amd",
              f"-Wl,-rpath,$ORIGIN"])
        # the same driver with the grid of the benchmark box compiled in (config.hpp's sizes
        # are compile-time, so the host layer is compiled into it rather than linked)
        _run([cxx, *flags, *inc, "-DMHD_NGRID1=256", "-DMHD_NGRID2=256", "-DMHD_NGRID3=256", "-o", KH_MAIN_256,
              main, *srcs, f"-L{PKG}", "-lmhd_amd", f"-Wl,-rpath,$ORIGIN"])
    if verbose:
        print(f"built {KH_LIB}")
    return KH_LIB


def build_calib(verbose: bool = True) -> str:
    """The calibration executable (stream bandwidth, FP64 issue rate)."""
    tmp = CALIB + ".tmp"
    _run([HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], "-o", tmp, os.path.join(PKG, "calib", "mhd_calib.hip")])
    os.replace(tmp, CALIB)
    if verbose:
        print(f"built {CALIB}")
    return CALIB


def build(verbose: bool = True) -> None:
    build_hip(verbose)
    build_host(verbose)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="build libmhd_amd.so (and the drop-in host layer) for gfx950")
    ap.add_argument("-q", "--quiet", action="store_true")
    ap.add_argument("-D", "--define", action="append", default=[],
                    help="compile switch NAME=VALUE for a build variant (MHD_FAST_DIV, MHD_FAST_SQRT, MHD_XY_DPP, "
                         "MHD_F64_MINWAVES)")
    ap.add_argument("--out", default=LIB, help="library path; a variant elsewhere skips the host layer")
    ap.add_argument("--calib", action="store_true", help="build only the mhd_calib measurement tool")
    a = ap.parse_args(argv)
    if a.calib:
        build_calib(not a.quiet)
    elif a.define or a.out != LIB:
        build_hip(not a.quiet, defines=a.define, out=os.path.abspath(a.out))
    else:
        build(not a.quiet)


if __name__ == "__main__":
    main()
