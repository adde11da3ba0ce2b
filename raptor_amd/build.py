"""Builds libraptor_rt.so (HIP, gfx950) in-tree with hipcc.

Flags that the parity contract depends on:
  -ffp-contract=off                        no FMA contraction of a*b+c
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE division and square root
  -fno-gpu-flush-denormals-to-zero         keep fp32 denormals
and, for the box:
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 math
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libraptor_rt.so")
SOURCES = ["api.hip", "bvh_build.hip", "trace.hip", "render.hip", "materials.hip", "postprocess.hip", "adjoint.hip", "invrender.hip", "batch.hip", "scene_gen.cpp"]
ARCH = os.environ.get("RAPTOR_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def build(verbose=False):
    """Compile every HIP source into raptor_amd/libraptor_rt.so."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libraptor_rt.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
