"""Builds the in-tree shared library i3elvis_amd/libi3g.so for gfx950.

hipcc compiles the HIP sources (device code for gfx950 only, packed-FP32
instructions disabled), gcc compiles the C host layer, hipcc links.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libi3g.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["i3g_stokes.hip", "i3g_markers.hip", "i3g_heat.hip", "i3g_net.hip", "i3g_gen.hip"]
C_SOURCES = ["i3g_host.c", "i3g_model.c", "i3g_io.c", "i3g_vtk.c"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for s in C_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run(["gcc", "-O2", "-fPIC", "-fopenmp", "-Wall", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-fopenmp", "-lm", "-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
