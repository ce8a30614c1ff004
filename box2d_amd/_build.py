"""Build the box2d_amd shared libraries in-tree with hipcc for gfx950.

Outputs (under box2d_amd/lib/):
  libbox2d_amd.so      the Box2D C API with the GPU step (host bookkeeping + HIP kernels)
  libb2amd_scenes.so   the public-API scene driver linked against libbox2d_amd.so

Numerics flags: contraction off (no fused multiply-add, matching the reference's
-ffp-contract=off), IEEE fp32 division and square root, no denormal flushing, and packed fp32
instructions disabled so every float operation is a separately rounded IEEE operation.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "box2d_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(ROOT, "_build", "obj")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("B2AMD_OFFLOAD_ARCH", "gfx950")

NUMERIC_FLAGS = [
    "-ffp-contract=off",
    "-fno-fast-math",
]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

COMMON = ["-O3", "-fPIC", "-fvisibility=hidden", "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function",
          "-Wno-missing-braces"]

HOST_SOURCES = ["world.cpp", "tree.cpp", "geometry.cpp", "capi.cpp", "batch.cpp", "host_threads.cpp", "snapshot.cpp"]
HIP_SOURCES = ["device.hip", "adjoint.hip", "broadphase.hip"]


def _hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build box2d_amd")
    return path


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return proc.stdout


def _compile(src, lang):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if lang == "hip":
        cmd = [_hipcc(), "-x", "hip", "-std=c++17", "-c", src, "-o", obj] + COMMON + NUMERIC_FLAGS + HIP_FLAGS
    else:
        cmd = [_hipcc(), "-x", "c++", "-std=c++17", "-c", src, "-o", obj] + COMMON + NUMERIC_FLAGS
    _run(cmd)
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = [(os.path.join(CSRC, s), "cpp") for s in HOST_SOURCES] + [(os.path.join(CSRC, s), "hip") for s in HIP_SOURCES]
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))

    lib = os.path.join(LIB, "libbox2d_amd.so")
    _run([_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib] + objs + ["-pthread"])

    scenes = os.path.join(LIB, "libb2amd_scenes.so")
    _run(["gcc", "-std=gnu17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I", INCLUDE,
          os.path.join(PKG, "scenes", "scenes.c"), "-o", scenes, "-L", LIB, "-lbox2d_amd",
          "-Wl,-rpath,$ORIGIN", "-lm"])
    if verbose:
        print(f"built {lib}\nbuilt {scenes}")
    return lib


if __name__ == "__main__":
    build()
