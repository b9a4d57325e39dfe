"""Build recipe: hipcc for the gfx950 library, g++ for the C++ front-end.

All outputs stay in-tree (spike_amd/_lib, spike_amd/_bin) so that they travel with the
source tree to the GPU machine.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
BIN_DIR = os.path.join(HERE, "_bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every float expression is evaluated as written (no FMA contraction),
# the precondition of bit-exact parity with the CPU oracle.  Packed FP32 instructions are
# disabled for gfx950.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip_library() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", f) for f in ("spike_hip.hip", "mlp.hip")]
    out = os.path.join(LIB_DIR, "libspike_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_all() -> dict:
    out = {"libspike_hip": build_hip_library()}
    return out


if __name__ == "__main__":
    print(build_all())
