"""Builds libtcym.so (the gfx950 HIP library behind include/tcym.h) in-tree.

    python -m topologicalcymsolver_amd.build        # or build() from Python

Each .hip translation unit is compiled with hipcc for gfx950 only, the
host-only C++ sources (configuration files) with the host C++ compiler, in
parallel, and all are linked into topologicalcymsolver_amd/libtcym.so.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtcym.so")
SOURCES = ["tcym_api.hip", "tcym_batch.hip", "tcym_kernels_su2.hip", "tcym_kernels_su3.hip"]
HOST_SOURCES = ["tcym_io.cpp"]
DEPS = ["tcym_backend.hpp", "tcym_device.hpp", "tcym_group.hpp", "tcym_kernels.inc", "tcym_reduce.hpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "c++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + CSRC]
# rocFFT: the inverse FFT of the quasi-particle initial condition
LIBS = ["-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay disabled; FP32 code is scalar VALU
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _compile(src, build_dir=None, extra=()):
    obj = os.path.join(build_dir or BUILD, os.path.splitext(src)[0] + ".o")
    if src in HOST_SOURCES:
        cmd = [CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC] + FLAGS + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed for %s:\n%s\n%s" % (cmd[0], src, r.stdout, r.stderr))
    return obj


def build(verbose=False, jobs=None, out=None, defines=()):
    """Compile and link; `out`/`defines` build a library elsewhere with extra
    compiler options."""
    lib = out or LIB
    bdir = os.path.join(os.path.dirname(lib), "_build") if out else BUILD
    os.makedirs(bdir, exist_ok=True)
    jobs = jobs or min(len(SOURCES) + len(HOST_SOURCES), os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, bdir, defines), SOURCES + HOST_SOURCES))
    tmp = lib + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + LIBS,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
