"""Build libgdtp.so in-tree: host C++ with ROCm clang (+OpenMP), HIP kernels for gfx950.

Usage: python -m gradientdomaintextureprocessing_amd.build   (or call build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgdtp.so")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["gdtp_mesh.cpp", "gdtp_hierarchy.cpp", "gdtp_assembly.cpp", "gdtp_boundary.cpp", "gdtp_vectorfield.cpp", "gdtp_ldlt.cpp",
                "gdtp_galerkin.cpp"]
HIP_SOURCES = ["gdtp_device.hip", "gdtp_mixed.hip", "gdtp_assembly_device.hip", "gdtp_capi.cpp"]

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-fopenmp",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wno-sign-compare", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    clang = os.path.join(ROCM, "llvm", "bin", "clang++")
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([clang] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    sys.path.insert(0, os.path.dirname(HERE))
    from gradientdomaintextureprocessing_amd._lib import source_hash
    build_id = ['-DGDTP_BUILD_ID="' + source_hash() + '"']
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        extra = build_id if src == "gdtp_capi.cpp" else []
        jobs.append([hipcc] + HIP_FLAGS + extra + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(BUILD, s + ".o") for s in HOST_SOURCES + HIP_SOURCES]
    tmp = OUT + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs +
         ["-Wl,-rpath," + os.path.join(ROCM, "llvm", "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
