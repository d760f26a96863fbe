"""Build script for the MI355X SCC library.

Compiles the gfx950 HIP sources and the host C++ sources into
``libpsc.so`` next to this file, and the ``scc`` command line (the
counterpart of src/scc.cpp) into ``bin/scc``.  Always recompiles.

    python parallel-strong-connectivity_amd/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpsc.so")
CLI = os.path.join(HERE, "bin", "scc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # keep packed fp32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wextra",
             f"-I{INCLUDE}", f"-I{CSRC}"]

HIP_SOURCES = ["scc_engine.hip", "tail.hip", "gen_device.hip", "batch.hip", "cc.hip",
               "le_list.hip", "noalloc.hip"]
CXX_SOURCES = ["graph_host.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _gomp_dir():
    out = subprocess.run([CXX, "-print-file-name=libgomp.so"], check=True,
                         capture_output=True, text=True).stdout.strip()
    return os.path.dirname(out)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        _run([CXX, *CXX_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    gomp = _gomp_dir()
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          f"-L{gomp}", "-lgomp", "-lpthread",
          f"-Wl,-rpath,{gomp}"])
    os.replace(tmp, LIB)
    # command-line driver, linked against the library by relative rpath
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    _run([CXX, "-O2", "-std=c++17", f"-I{INCLUDE}",
          os.path.join(CSRC, "scc_main.cpp"), "-o", CLI,
          f"-L{HERE}", "-lpsc", "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    try:
        build()
    except subprocess.CalledProcessError as e:
        sys.exit(e.returncode)
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        sys.exit("hipcc not found")
