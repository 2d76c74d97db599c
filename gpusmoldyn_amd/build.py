"""Builds libgpusmoldyn.so in-tree with hipcc for gfx950.

Usage: python -m gpusmoldyn_amd.build   (or gpusmoldyn_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgpusmoldyn.so")

# -ffp-contract=off: every fused multiply-add is an explicit fmaf so the CPU
# oracle can reproduce results bit for bit.  Packed FP32 instructions are
# disabled for every HIP translation unit.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]

SOURCES = ["kernels.hip", "legacy.hip", "copy_probe.hip", "engine.cpp", "setup.cpp", "rdf.cpp", "rdf_gpu.hip", "capi.cpp", "particle_system.cpp", "sfmt.cpp", "noalloc.hip"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs + ["-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        l for l in (res.stdout + res.stderr).splitlines()
        if "packed-fp32-ops" not in l and "not a recognized feature" not in l
    )
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


DRIVER = os.path.join(HERE, "smd_particles")


def build_driver(verbose=False):
    """The headless particles.cpp-style driver, linked against the library."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", inc, os.path.join(CSRC, "particles_main.cpp"),
           "-L", HERE, "-lgpusmoldyn", "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("driver build failed:\n" + res.stdout + res.stderr)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
