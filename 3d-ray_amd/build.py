"""Builds lib3dray_amd.so (HIP kernels + C ABI) for gfx950 with hipcc.

Flags that matter for parity: -ffp-contract=off (no FMA contraction, so the
fp32 operation order is the source order), IEEE-correct fp32 division and
square root, denormals kept.  -Xclang -target-feature -Xclang -packed-fp32-ops
keeps the compiler from emitting packed fp32 arithmetic.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib3dray_amd.so")
SOURCES = ["scene_build.hip", "render.hip", "film_io.hip", "denoise.hip"]
HEADERS = ["rt_common.h", "rt_internal.h", "rt_scene.h", "traverse.h", "shade.h", "material.h", "adjoint.h", "photon.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-pthread",
]


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-x", "hip"] + [os.path.join(CSRC, s) for s in SOURCES] + ["-lz", "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building lib3dray_amd.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
