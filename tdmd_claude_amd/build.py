"""Build the gfx950 shared library ``tdmd_claude_amd/lib/libtdmd_hip.so``.

Every translation unit under ``csrc/`` is compiled by hipcc as HIP for
gfx950 only, then linked into one shared library that exports the C-ABI
declared in ``include/tdmd_hip.h``. Run ``python -m tdmd_claude_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libtdmd_hip.so"

SOURCES = [
    "runtime.cpp",
    "scan.cpp",
    "cells.cpp",
    "nlist_csr.cpp",
    "tables.cpp",
    "forces_csr.cpp",
    "integrate.cpp",
    "scheduler.cpp",
    "fm.cpp",
    "pipeline.cpp",
]

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROCM = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _compile(src: str) -> pathlib.Path:
    obj = BUILD / (pathlib.Path(src).stem + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


CPP_TEST = ROOT / "tests" / "cpp" / "dropin_test.cpp"
CPP_IO_TEST = ROOT / "tests" / "cpp" / "io_test.cpp"
CPP_BIN = ROOT / "tests" / "cpp" / "_bin"


def build_cpp_tests(verbose: bool = True) -> list[pathlib.Path]:
    """Compile the C++ drop-in test program once per precision macro. It is
    host code over ``include/tdmd/`` and links the library by rpath."""
    CPP_BIN.mkdir(parents=True, exist_ok=True)
    out = []
    for prec in ("MIXED", "FP64"):
        exe = CPP_BIN / f"dropin_test_{prec.lower()}"
        # hip_runtime_api.h for host compilers: the AMD platform macro and
        # the ROCm include directory (what hipcc adds for host code).
        cmd = [CXX, "-std=c++20", "-O2", "-Wall", f"-DTDMD_PRECISION_{prec}",
               "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}",
               str(CPP_TEST), f"-L{LIBDIR}", "-ltdmd_hip",
               "-Wl,-rpath,$ORIGIN/../../../tdmd_claude_amd/lib", "-o", str(exe)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"C++ test build failed:\n{' '.join(cmd)}\n{res.stdout}\n"
                               f"{res.stderr}")
        out.append(exe)
        if verbose:
            print(f"built {exe}")
        # host-only io checks: no device code, no library
        exe = CPP_BIN / f"io_test_{prec.lower()}"
        cmd = [CXX, "-std=c++20", "-O2", "-Wall", f"-DTDMD_PRECISION_{prec}",
               "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}",
               str(CPP_IO_TEST), "-o", str(exe)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"C++ io test build failed:\n{' '.join(cmd)}\n{res.stdout}\n"
                               f"{res.stderr}")
        out.append(exe)
        if verbose:
            print(f"built {exe}")
    return out


def build(verbose: bool = True) -> pathlib.Path:
    """Compile all sources and link the shared library. Always rebuilds."""
    if not pathlib.Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cpp_tests(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
