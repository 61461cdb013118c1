"""Builds libbxstep.so (HIP for gfx950) in-tree: bounce_amd/lib/libbxstep.so.

Floating-point flags are part of the product's contract (bit-exact binary32 against the
reference): no FMA contraction, IEEE division and square root, no denormal flushing, and packed
FP32 instructions disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libbxstep.so")
SOURCES = ["bx_step.hip", "bx_api.hip"]
HEADERS = ["bx_math.h", "bx_types.h", "bx_collide.h", "bx_convex.h", "bx_triangle.h", "bx_mesh.h", "bx_joint.h", "bx_solver.h", "bx_shapes.h", "bx_islands.h",
           "bx_solve_kernels.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BX_OFFLOAD_ARCH", "gfx950")

FP_FLAGS = [
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def command(out=LIB, extra=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    return ([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-I", os.path.join(REPO, "include"), "-Wno-unused-result"] + FP_FLAGS + list(extra)
            + ["-o", out] + srcs)


CAL_LIB = os.path.join(LIB_DIR, "libbxcal.so")
FACADE_LIB = os.path.join(LIB_DIR, "libbounce_amd.so")
HELLO_BIN = os.path.join(LIB_DIR, "bx_hello_world")
MESH_BIN = os.path.join(LIB_DIR, "bx_mesh_ground")
SPAWN_BIN = os.path.join(LIB_DIR, "bx_spawn_apply")
LIB_DIR_RPATH = "$ORIGIN"


def build(verbose=True):
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = [(LIB, command(LIB + ".tmp")),
            # measurement support: plain copy kernel for calibrating the profiler's byte counters
            (CAL_LIB, [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared"] + FP_FLAGS
             + ["-o", CAL_LIB + ".tmp", os.path.join(CSRC, "bx_calibrate.hip")])]
    # the drop-in C++ API (bounce/bounce.h) over the C ABI, and a program built against it
    host = ["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
            "-I", os.path.join(PKG, "include"), "-I", os.path.join(REPO, "include")]
    jobs.append((FACADE_LIB, host + ["-shared", "-o", FACADE_LIB + ".tmp", os.path.join(PKG, "facade", "world.cpp"),
                                     "-L", LIB_DIR, "-lbxstep", "-Wl,-rpath,$ORIGIN"]))
    jobs.append((HELLO_BIN, host + ["-o", HELLO_BIN + ".tmp", os.path.join(REPO, "tests", "drop_in", "hello_world.cpp"),
                                    "-L", LIB_DIR, "-lbounce_amd", "-lbxstep", "-Wl,-rpath," + LIB_DIR_RPATH]))
    jobs.append((MESH_BIN, host + ["-o", MESH_BIN + ".tmp", os.path.join(REPO, "tests", "drop_in", "mesh_ground.cpp"),
                                   "-L", LIB_DIR, "-lbounce_amd", "-lbxstep", "-Wl,-rpath," + LIB_DIR_RPATH]))
    jobs.append((SPAWN_BIN, host + ["-o", SPAWN_BIN + ".tmp", os.path.join(REPO, "tests", "drop_in", "spawn_apply.cpp"),
                                    "-L", LIB_DIR, "-lbounce_amd", "-lbxstep", "-Wl,-rpath," + LIB_DIR_RPATH]))
    for out, cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
