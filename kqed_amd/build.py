"""Build the HIP library kqed_amd/libkqed_hip.so for gfx950 (and nothing else).

    python -m kqed_amd.build            # or kqed_amd.build.build()

hipcc cross-compiles without a GPU.  Objects go to kqed_amd/_build/, the library next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkqed_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Every HIP translation unit: gfx950 only, no packed-FP32 instruction selection, and no implicit
# contraction (every fused multiply-add in the kernels is written out).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-ffp-contract=off"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for c in ("kqed_tables.c", "kqed_assembly.c", "kqed_pert_host.c"):
        o = os.path.join(OUT, c + ".o")
        _run(["gcc", *C_FLAGS, "-c", os.path.join(SRC, c), "-o", o])
        objs.append(o)
    for h in ("kqed_kernels.hip", "kqed_surrogate.hip", "kqed_combine.hip", "kqed_pert.hip", "kqed_api.hip"):
        o = os.path.join(OUT, h + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(SRC, h), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lm"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
