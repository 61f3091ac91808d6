"""Build libmlh_hip.so (HIP kernels + C ABI) for gfx950, in-tree.

    python -m meshlesshydro_amd.build

hipcc compiles every .hip source of csrc/ into one shared library with the
C ABI of include/mlh.h.  Packed FP32 instructions are disabled for device
code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmlh_hip.so")

SOURCES = ["api.hip", "cells.hip", "hydro.hip", "hydro_fp32.hip", "mlp.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    # FP32 divide / sqrt as the hardware reciprocal / square root (~1 ulp);
    # FP64 is unaffected (FP32 arithmetic only runs under stated tolerances)
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # FP32 denormals flushed to zero: the FP32 divide and square root then
    # need no range scaling around the hardware instructions.  FP64 keeps
    # IEEE denormals (its code is unchanged by this flag).
    "-fgpu-flush-denormals-to-zero",
    "-Wno-unused-result",
]


def build(verbose=False, out=OUT, extra=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [hipcc] + FLAGS + (extra or []) + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmlh_hip.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
