#!/usr/bin/env python3
"""Build the MI355X engine and its hosts in-tree.

Outputs (git-ignored, shipped with the tree):
  lib/libiris.so        HIP engine for gfx950 (C ABI of include/iris.h)
  lib/libiris_host.so   Fortran host set-up (inputFDTD.dat -> iris_params)
  bin/irisfdtd          drop-in Fortran program (reference program flow)

Usage: python build.py [--only engine|host|all]
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("IRIS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # the FP64 path reproduces the reference's operation order: no contraction
    "-ffp-contract=off",
    # packed FP32 forms are avoided on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    src = os.path.join(HERE, "csrc", "iris_engine.hip")
    out = os.path.join(LIB, "libiris.so")
    tmp = out + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    fdir = os.path.join(HERE, "fortran")
    # amdflang accepts -J for the module directory
    fflags = ["-O2", "-fPIC", "-J", OBJ]
    objs = []
    for name in ["iris_c_api", "iris_sdl", "iris_setup"]:
        o = os.path.join(OBJ, name + ".o")
        run([FC] + fflags + ["-c", os.path.join(fdir, name + ".f90"), "-o", o])
        objs.append(o)
    host = os.path.join(LIB, "libiris_host.so")
    run([FC, "-shared", "-fPIC", "-o", host + ".tmp"] + objs)
    os.replace(host + ".tmp", host)
    mo = os.path.join(OBJ, "iris_main.o")
    run([FC] + fflags + ["-c", os.path.join(fdir, "iris_main.f90"), "-o", mo])
    exe = os.path.join(BIN, "irisfdtd")
    run([FC, "-o", exe + ".tmp", mo, "-L", LIB, "-liris_host", "-liris",
         "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(exe + ".tmp", exe)
    return host, exe


def main(argv):
    what = "all"
    if "--only" in argv:
        what = argv[argv.index("--only") + 1]
    if what in ("engine", "all"):
        build_engine()
    if what in ("host", "all"):
        build_host()
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv[1:])
