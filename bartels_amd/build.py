"""Builds the in-tree gfx950 shared library ``bartels_amd/libbartels_hip.so`` with hipcc.

Run ``python -m bartels_amd.build`` (or ``__graft_entry__.build()``).  The library is the product:
hand-written HIP kernels behind the C ABI declared in ``include/bartels_hip.h``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libbartels_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "bartels_hip.hip"), os.path.join(HERE, "csrc", "io.cpp")]
HEADERS = [os.path.join(HERE, "csrc", "elements.h"), os.path.join(HERE, "csrc", "newton_impl.h"),
           os.path.join(HERE, "csrc", "tri_impl.h"),
           os.path.join(ROOT, "include", "bartels_hip.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions (v_pk_*_f32) stay off: some forms have given wrong results on MI355X
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = True) -> str:
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include")] + SOURCES + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
