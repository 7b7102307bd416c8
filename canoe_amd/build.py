"""Build libcanoe_gpu.so (HIP kernels for gfx950 + the C-ABI) in-tree.

    python -m canoe_amd.build            # or: python canoe_amd/build.py

Every device kernel is compiled with -ffp-contract=off (no fused
multiply-add: the FP64 path reproduces a CPU evaluation bit for bit) and
with packed FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcanoe_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CANOE_GPU_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-value", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]

SOURCES_HIP = ["capi.hip", "run.hip", "mlp.hip"]
SOURCES_CPP = ["pgen.cpp", "netcdf.cpp", "sphere_geom.cpp"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in SOURCES_HIP:
        o = os.path.join(objdir, s + ".o")
        run([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, s), "-o", o]
            + COMMON + DEVICE)
        objs.append(o)
    for s in SOURCES_CPP:
        o = os.path.join(objdir, s + ".o")
        run(["g++", "-c", os.path.join(CSRC, s), "-o", o] + COMMON)
        objs.append(o)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH,
                                                "-fPIC"])
    os.replace(tmp, LIB)
    # C++ host driver over the C-ABI (the reference's main loop)
    exe = os.path.join(LIBDIR, "canoe_step")
    run(["g++", "-std=c++17", "-O2", "-Wall",
         "-I" + os.path.join(ROOT, "include"),
         os.path.join(CSRC, "canoe_step.cpp"), "-o", exe + ".tmp",
         LIB, "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
