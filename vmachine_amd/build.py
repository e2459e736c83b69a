"""Builds libvmc_hip.so in-tree (vmachine_amd/_lib/) with hipcc for gfx950, and
the `vmachine` executable (host C++ over the library's C ABI) beside it.

Usage: python -m vmachine_amd.build   (or __graft_entry__.build())
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libvmc_hip.so")
OBJDIR = os.path.join(PKG, "_lib", "obj")
EXE = os.path.join(LIBDIR, "vmachine")
HOST_LIB = os.path.join(LIBDIR, "libvmc_host.so")
HOST_SOURCES = ["config.cpp", "runner.cpp", "cli.cpp"]
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

ARCH = os.environ.get("VMC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# packed FP32 instructions are kept out of every code object.  -fno-signed-zeros
# lets the compiler drop additions of a zero (a sum started at 0, fma(a, b, 0));
# the sign of a zero reaches no result of the engine.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fno-signed-zeros",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

# the walk kernels are split by precision (and the closed shells past N = 56
# into units of their own) so that they compile in parallel
UNITS = [
    ("vmc_walk_f32.o", "vmc_walk.hip", ["-DVMC_WALK_PRECISION=32"]),
    ("vmc_walk_f64.o", "vmc_walk.hip", ["-DVMC_WALK_PRECISION=64"]),
    ("vmc_walk_common.o", "vmc_walk.hip", ["-DVMC_WALK_PRECISION=0"]),
    ("vmc_walk_large_f32.o", "vmc_walk.hip", ["-DVMC_WALK_PRECISION=32", "-DVMC_WALK_LARGE"]),
    ("vmc_walk_large_f64.o", "vmc_walk.hip", ["-DVMC_WALK_PRECISION=64", "-DVMC_WALK_LARGE"]),
    ("vmc_learn.o", "vmc_learn.hip", []),
    ("vmc_batch.o", "vmc_batch.hip", []),
    ("vmc_optim.o", "vmc_optim.hip", []),
    ("vmc_engine.o", "vmc_engine.cpp", []),
]


def _compile(unit):
    obj, src, extra = unit
    out = os.path.join(OBJDIR, obj)
    cmd = [HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    msgs = [ln for ln in r.stdout.splitlines() if "unknown target feature" not in ln
            and "packed-fp32-ops" not in ln and ln.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, "\n".join(msgs)))
    return out


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_exe()
    if verbose:
        print("built", LIB, "and", EXE)
    return LIB


def _host(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n" + r.stdout)


def build_exe():
    """The host layer is plain C++17 (g++): libvmc_host.so reaches the GPU only
    through libvmc_hip.so's C ABI, and `vmachine` is a one-line main over it."""
    inc = "-I" + os.path.join(ROOT, "include")
    srcs = [os.path.join(CSRC, "host", f) for f in HOST_SOURCES]
    tmp = HOST_LIB + ".tmp"
    _host([CXX, "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", inc, "-o", tmp] + srcs
          + ["-L" + LIBDIR, "-lvmc_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    tmp = EXE + ".tmp"
    _host([CXX, "-O2", "-std=c++17", "-Wall", inc, "-o", tmp,
           os.path.join(CSRC, "host", "main.cpp"), "-L" + LIBDIR, "-lvmc_host", "-lvmc_hip",
           "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, EXE)
    return EXE


if __name__ == "__main__":
    build()
    sys.exit(0)
