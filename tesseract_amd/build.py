"""Builds the native library in-tree: tesseract_amd/libtesseract.so.

Host code (program model, program writer and .bench loader, specs, random programs, reports) is
compiled with gcc / g++; the engine (C++ over the HIP runtime and hipRTC, plus the hand-written
gfx950 kernels) with hipcc for gfx950.
Run as `python -m tesseract_amd.build`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libtesseract.so")

HOST_C = ["program", "tester", "benes", "report"]
HOST_CPP = ["spec", "circuit"]
ENGINE_CPP = ["lower", "codegen", "tape", "tape_kernel", "engine", "calibrate", "batch", "serve", "tester1",
              "clear", "noalloc"]
ARCH = "gfx950"
# Packed FP32 stays off in every HIP compile.
NO_PACKED = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    cmds, objs = [], []
    for name in HOST_C:
        src = os.path.join(CSRC, "host", name + ".c")
        obj = os.path.join(OBJ, name + ".o")
        cmds.append(["gcc", "-std=gnu11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-maybe-uninitialized",
                     "-I" + INCLUDE, "-c", src, "-o", obj])
        objs.append(obj)
    for name in HOST_CPP:
        src = os.path.join(CSRC, "host", name + ".cpp")
        obj = os.path.join(OBJ, name + ".o")
        cmds.append(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-I" + INCLUDE, "-c", src,
                     "-o", obj])
        objs.append(obj)
    for name in ENGINE_CPP:
        src = os.path.join(CSRC, "engine", name + ".cpp")
        if not os.path.exists(src):
            src = os.path.join(CSRC, "engine", name + ".hip")
        if not os.path.exists(src):
            continue
        obj = os.path.join(OBJ, name + ".o")
        cmd = ["hipcc", "-std=c++17", "-O2", "-fPIC", "-I" + INCLUDE, "-Wno-unused-command-line-argument",
               "--offload-arch=" + ARCH] + NO_PACKED
        if src.endswith(".hip"):
            cmd += ["-O3"]
        cmds.append(cmd + ["-c", src, "-o", obj])
        objs.append(obj)
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs +
         ["-lhiprtc", "-lamdhip64", "-lpthread"], verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
