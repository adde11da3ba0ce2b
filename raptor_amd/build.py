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
SOURCES = ["api.hip", "bvh_build.hip", "trace.hip", "render.hip", "materials.hip", "postprocess.hip", "adjoint.hip", "invrender.hip", "batch.hip", "scene_gen.cpp", "ingest.cpp", "jpeg.hip", "glare.hip", "aac_build.hip"]
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


def build(verbose=False, jobs=None):
    """Compile every HIP source into raptor_amd/libraptor_rt.so: one object
    per source, compiled in parallel under _build/, then one link."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(os.path.dirname(HERE), "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + compile_flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        return obj, res

    jobs = jobs or min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(compile_one, SOURCES))
    failed = [res for _, res in results if res.returncode != 0]
    for _, res in results:
        # hipcc's host pass does not know the packed-fp32 feature and says so.
        text = "".join(line for line in (res.stdout + res.stderr).splitlines(True)
                       if "packed-fp32-ops' is not a recognized feature" not in line)
        if text.strip() and (verbose or res.returncode != 0):
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("hipcc failed building libraptor_rt.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [obj for obj, _ in results]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libraptor_rt.so")
    os.replace(tmp, LIB)
    return LIB

if __name__ == "__main__":
    print(build(verbose=True))
