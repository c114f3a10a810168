"""Build libdlgpu.so (HIP, gfx950) in-tree.

    python -m dynlattice_amd.build

Compiles the kernels (csrc/*.hip) and the C ABI host code (csrc/*.cpp) with
hipcc for gfx950 into dynlattice_amd/libdlgpu.so, and the host
instantiation of the device math used by the CPU unit tests
(csrc/dl_hostcheck.cpp) into dynlattice_amd/libdlhost.so.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdlgpu.so")
HOSTLIB = os.path.join(HERE, "libdlhost.so")
HOSTCHECK = "dl_hostcheck.cpp"
ARCH = os.environ.get("DL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions off (FP64 path; keeps the compiler from
    # emitting v_pk_* f32 forms anywhere in the library)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument", "-Wno-unused-value",
]


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"),
              shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src, bdir, verbose):
    o = os.path.join(bdir, os.path.basename(src) + ".o")
    cmd = [hipcc()] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) + ["-c", src, "-o", o]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return o


def _link(objs, out, verbose):
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + \
        sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    jobs = max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, bdir, verbose), srcs))
    host = [o for s, o in zip(srcs, objs) if os.path.basename(s) == HOSTCHECK]
    _link([o for o in objs if o not in host], LIB, verbose)
    _link(host, HOSTLIB, verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
