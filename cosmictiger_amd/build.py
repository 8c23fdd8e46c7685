"""Builds libcosmictiger_amd.so (HIP for gfx950 + host C++) in-tree.

The library is the product: the C-ABI declared in include/cosmictiger_amd.h.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcosmictiger_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed FP32 instructions are disabled (see the pool notes on
# v_pk_* FP32 forms) and host code is built for any x86-64 with AVX2/FMA.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
HOST_ARCH = ["-mavx2", "-mfma"]
HIP_SOURCES = ["ct_kernels.hip", "ct_direct.hip", "ct_treegpu.hip", "ct_rungs.hip", "ct_io.hip"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def generate():
    run([sys.executable, os.path.join(HERE, "codegen", "gen_expansions.py"), os.path.join(CSRC, "ct_expansions.hpp")])


def build(regenerate=True):
    if regenerate:
        generate()
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    tree_o = os.path.join(build_dir, "ct_tree.o")
    # no FMA contraction in the host tree build: the device build
    # (ct_treegpu.hip) repeats its arithmetic operation for operation
    run([HIPCC] + COMMON + HOST_ARCH + ["-ffp-contract=off", "-x", "c++", "-c", os.path.join(CSRC, "ct_tree.cpp"), "-o",
                                        tree_o])
    objs.append(tree_o)
    for src in HIP_SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        run([HIPCC] + COMMON + DEVICE_FLAGS + ["-Xarch_host", "-mavx2", "-Xarch_host", "-mfma", "-c",
                                               os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC] + DEVICE_FLAGS + ["-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
