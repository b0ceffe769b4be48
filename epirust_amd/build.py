"""Builds libepigpu.so (gfx950) and the epirust-gpu executable in-tree with hipcc.

    python -m epirust_amd.build        # or: python epirust_amd/build.py

Objects go to _build/ at the repository root; the shared library and the executable are
written next to this file so they travel with the source tree.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(ROOT, "_build", "epigpu")
LIB = os.path.join(PKG, "libepigpu.so")
EXE = os.path.join(PKG, "epirust-gpu")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-I" + SRC, "-I" + os.path.join(ROOT, "include")]
HOST = ["-I" + os.path.join(ROCM, "include"), "-D__HIP_PLATFORM_AMD__"]
# Packed FP32 instructions are kept out of every code object built here.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HOST_SOURCES = ["config.cpp", "population.cpp", "engine.cpp", "cli.cpp", "mlp_engine.cpp"]
DEVICE_SOURCES = ["kernels.hip", "init.hip", "stream.hip", "mlp.hip"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HOST_SOURCES:
        jobs.append([CXX] + COMMON + HOST + ["-c", os.path.join(SRC, s), "-o", os.path.join(OBJ, s + ".o")])
    for s in DEVICE_SOURCES:
        jobs.append([HIPCC] + COMMON + DEVICE + ["-c", os.path.join(SRC, s), "-o", os.path.join(OBJ, s + ".o")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [os.path.join(OBJ, s + ".o") for s in HOST_SOURCES + DEVICE_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    main_obj = os.path.join(OBJ, "main.cpp.o")
    _run([CXX] + COMMON + HOST + ["-c", os.path.join(SRC, "main.cpp"), "-o", main_obj])
    _run([HIPCC, "-o", EXE, main_obj, "-L" + PKG, "-lepigpu", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
