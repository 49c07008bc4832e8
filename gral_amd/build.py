"""Builds the HIP extension libgral_pf.so in-tree for gfx950 with hipcc.

Flags: -ffp-contract=off keeps the reference's unfused FP32 arithmetic (no FMA
contraction); division and square root stay correctly rounded (the HIP
default); packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgral_pf.so")
SOURCES = [os.path.join(HERE, "csrc", "pf_kernels.hip"), os.path.join(HERE, "csrc", "pf_api.cpp"),
           os.path.join(HERE, "csrc", "pf_zip.cpp"), os.path.join(HERE, "csrc", "pf_init.hip"),
           os.path.join(HERE, "csrc", "pf_adjoint.hip")]
LIBS = ["-lz"]  # zlib: deflate of the flow-field files' zip container
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def build(verbose: bool = False) -> str:
    cmd = [hipcc()] + FLAGS + ["-x", "hip"] + SOURCES + LIBS + ["-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgral_pf.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
