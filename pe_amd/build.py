"""Builds the HIP library ``pe_amd/_lib/libpehip.so`` for gfx950 (in-tree).

Every translation unit is compiled with hipcc for gfx950 only, with FP
contraction and fast-math off (the parity kernels keep pe's floating-point
operation order) and without packed-FP32 instruction selection.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libpehip.so")

SOURCES = ["k_bodies.hip", "k_broadphase.hip", "k_solver.hip", "k_islands.hip", "k_batch.hip", "k_adjoint.hip", "context.hip",
           "batch.hip", "adjoint.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]


def compile_one(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose=False):
    """Compiles all sources and links the shared library; returns its path."""
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(len(srcs), 6)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_examples(verbose)
    return LIB


def build_examples(verbose=False):
    """C++ programs on the drop-in API (include/pe_amd/pe.hpp), linked against the library."""
    exdir = os.path.join(HERE, "examples")
    for src in sorted(os.listdir(exdir)) if os.path.isdir(exdir) else []:
        if not src.endswith(".cpp"):
            continue
        exe = os.path.join(OUT, src[:-4])
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), os.path.join(exdir, src),
               "-o", exe, "-L", OUT, "-lpehip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("example build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
        if verbose:
            print("built", exe)


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)
    shutil.rmtree(OUT, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
