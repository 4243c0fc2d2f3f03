"""Builds libhodlrnd.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled with -ffp-contract=off (entry
evaluation must round like the FP64 CPU restatement; fused multiply-adds are
written out explicitly) and without packed-FP32 instructions.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhodlrnd.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["tree.cpp", "op.cpp", "opio.cpp", "vecio.cpp", "solver.cpp", "batch_solver.cpp", "server.cpp", "capi.cpp", "aca.hip",
           "matvec.hip", "blas.hip", "skeleton.hip", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-6000:])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


def build_examples(verbose=True):
    """The reference-shaped C++ API examples: host compiler + the C ABI, and
    device_kernels (hipcc) whose kernel classes' Kernel_Fun runs on the GPU."""
    out_dir = os.path.join(HERE, "examples", "build")
    os.makedirs(out_dir, exist_ok=True)
    root = os.path.dirname(HERE)
    outs = []
    for name in ("main_4d", "user_kernels"):
        out = os.path.join(out_dir, name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(root, "include"),
               os.path.join(HERE, "examples", name + ".cpp"), "-o", out, "-L" + HERE, "-lhodlrnd",
               "-Wl,-rpath,$ORIGIN/../.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("example build failed:\n" + r.stderr[-4000:])
        if verbose:
            print("built", out)
        outs.append(out)
    # kernel classes compiled as host and device code: hipcc, with the same
    # floating-point flags as the library
    out = os.path.join(out_dir, "device_kernels")
    cmd = [HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-Xclang",
           "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-I" + os.path.join(root, "include"),
           os.path.join(HERE, "examples", "device_kernels.cpp"), "-o", out, "-L" + HERE, "-lhodlrnd",
           "-Wl,-rpath,$ORIGIN/../.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("example build failed:\n" + r.stderr[-4000:])
    if verbose:
        print("built", out)
    outs.append(out)
    return outs


if __name__ == "__main__":
    build()
    build_examples()
    sys.exit(0)
