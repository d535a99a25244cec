"""Builds the HIP library in-tree for gfx950 (no JIT cache, no pip install).

    python -m lajolla_public_amd.build

produces lajolla_public_amd/liblajolla_amd.so from csrc/.  hipcc cross-compiles
here without a GPU.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liblajolla_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["ljx_lib.hip", "ljx_guide_net.hip", "ljx_bvh_gpu.hip"]
CXX_SOURCES = ["ljx_bvh.cpp", "ljx_parse.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-result"]
# Packed FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in CXX_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        run([HIPCC, "-x", "hip", *COMMON, *DEVICE, "-pthread", "-I" + CSRC, "-c", os.path.join(CSRC, src),
             "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        run([HIPCC, *COMMON, *DEVICE, "-I" + CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-pthread", "-o", tmp, *objs, "-lz"])
    os.replace(tmp, OUT)
    # the command-line renderer, a host program over the C++ API
    cli = os.path.join(HERE, "lajolla_amd")
    run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"), os.path.join(HERE, "host", "main.cpp"),
         "-L" + HERE, "-llajolla_amd", "-Wl,-rpath,$ORIGIN", "-o", cli + ".tmp"])
    os.replace(cli + ".tmp", cli)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
