"""Builds libepigpu.so (gfx950) and the epirust-gpu executable in-tree with hipcc.

    python -m epirust_amd.build        # or: python epirust_amd/build.py

Objects go to _build/ at the repository root; the shared library and the executable are
written next to this file so they travel with the source tree.

A library with another disease handler (csrc/disease.hpp describes the handler type) is
built from the default objects plus the three sources that depend on the handler:

    python -m epirust_amd.build --disease-handler path/to/my_disease.hpp:epi::MyDisease --out libepigpu_my.so

and selected at run time with EPIRUST_GPU_LIB=libepigpu_my.so.
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
# the sources that call the disease handler (csrc/disease.hpp)
HANDLER_SOURCES = ["engine.cpp", "kernels.hip", "init.hip"]


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


def build_handler_variant(header, handler, out, verbose=False):
    """libepigpu.so with the disease handler `handler` (a qualified type name) declared in
    `header` compiled in place of DefaultDisease; written to `out`. Needs build() first."""
    header = os.path.abspath(header)
    if not os.path.exists(header):
        raise RuntimeError("disease handler header not found: %s" % header)
    tag = "".join(c if c.isalnum() else "_" for c in os.path.splitext(os.path.basename(out))[0])
    obj_dir = os.path.join(OBJ, "variant_" + tag)
    os.makedirs(obj_dir, exist_ok=True)
    defs = ['-DEPI_DISEASE_HEADER="%s"' % header, "-DEPI_DISEASE_HANDLER=%s" % handler]
    jobs = []
    for s in HANDLER_SOURCES:
        tool, flags = (HIPCC, DEVICE) if s.endswith(".hip") else (CXX, HOST)
        jobs.append([tool] + COMMON + flags + defs + ["-c", os.path.join(SRC, s), "-o", os.path.join(obj_dir, s + ".o")])
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for o in ex.map(_run, jobs):
            if verbose and o.strip():
                print(o)
    objs = [os.path.join(obj_dir if s in HANDLER_SOURCES else OBJ, s + ".o") for s in HOST_SOURCES + DEVICE_SOURCES]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError("build the default library first (missing %s)" % ", ".join(missing))
    out = os.path.abspath(out)
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, out)
    return out


def main(argv):
    verbose = "-v" in argv
    lib = build(verbose=verbose)
    if "--disease-handler" in argv:
        header, _, handler = argv[argv.index("--disease-handler") + 1].partition(":")
        out = argv[argv.index("--out") + 1] if "--out" in argv else os.path.join(PKG, "libepigpu_variant.so")
        if not header or not handler:
            raise SystemExit("--disease-handler HEADER:TYPE")
        lib = build_handler_variant(header, handler, out, verbose=verbose)
    print(lib)


if __name__ == "__main__":
    main(sys.argv[1:])
