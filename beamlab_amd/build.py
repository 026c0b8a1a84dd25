"""Builds the gfx950 shared library of the package (beamlab_amd/libbeamlab_gpu.so) with hipcc, in-tree.

Floating-point flags are part of the contract: no fused multiply-add contraction, IEEE fp32 division and square
root, no fast-math, denormals kept - the step reproduces the reference CPU arithmetic bit for bit. Packed FP32
instructions are switched off (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbeamlab_gpu.so")
SOURCES = ["bl_step.hip", "bl_contacts.hip", "bl_capi.hip", "bl_fem.hip", "bl_gnn.hip"]
HEADERS = ["bl_devmath.h", "bl_world.h", "bl_contacts.h", "bl_fem_api.h", "bl_islands.h"]
ARCH = os.environ.get("BL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unknown-pragmas",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build_library(verbose=False):
    """Compiles every HIP source into libbeamlab_gpu.so (always recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}"] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libbeamlab_gpu.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
