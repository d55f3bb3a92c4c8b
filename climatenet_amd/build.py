"""Builds libpdfc.so (HIP, gfx950) in-tree: climatenet_amd/lib/libpdfc.so.

    python -m climatenet_amd.build

hipcc compiles the kernels and the host layer into one shared library. Flags:
  -ffp-contract=off   contact predicates must not be fused (bit-exact contact sets
                      against the CPU oracle); hot loops use explicit fma.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libpdfc.so")
SOURCES = ["device.hip", "selfcollide.hip", "mlp.hip", "sim.cpp", "model.cpp", "factor.cpp", "selfcontact.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PDFC_OFFLOAD_ARCH", "gfx950")


def flags():
    return [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
        "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    ]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cpp_api(verbose)
    return LIB


CPP_SOURCES = ["json.cpp", "scene_building.cpp", "simulation.cpp"]


def build_cpp_api(verbose=False):
    """The C++ host interface (reference API mirror) over the C ABI, and the driver."""
    cpp = os.path.join(HERE, "cpp")
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + cpp]
    rp = "-Wl,-rpath,$ORIGIN"
    api = os.path.join(LIBDIR, "libpdfc_cpp.so")
    cmds = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + inc + [os.path.join(cpp, s) for s in CPP_SOURCES]
        + ["-L" + LIBDIR, "-lpdfc", rp, "-o", api],
        ["g++", "-O2", "-std=c++17", "-Wall"] + inc + [os.path.join(cpp, "main.cpp"), "-L" + LIBDIR, "-lpdfc_cpp",
                                                      "-lpdfc", rp, "-o", os.path.join(LIBDIR, "pdfc_main")],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
