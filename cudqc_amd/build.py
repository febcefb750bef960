"""Builds the in-tree native library and CLI for gfx950 with hipcc.

Outputs (git-ignored, shipped with the tree):
  cudqc_amd/lib/libcudqc_amd.so  - C-ABI of include/dqc.h (kernels + host path)
  cudqc_amd/bin/cudqc            - reference-compatible command line
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
LIB = os.path.join(LIB_DIR, "libcudqc_amd.so")
CLI = os.path.join(BIN_DIR, "cudqc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-fp32 instructions are disabled for all device code built here
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wno-unused-value", "-Wno-unused-result", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include")]
SOURCES = ["engine.hip", "costmodel.hip", "graph.cpp", "rmnonmax.cpp", "search.cpp", "gen.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    objs = []
    obj_dir = os.path.join(HERE, "lib", "obj")
    os.makedirs(obj_dir, exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(obj_dir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + lang + ["-c", path, "-o", obj]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fopenmp", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "cli.cpp"),
          "-L", LIB_DIR, "-lcudqc_amd", "-Wl,-rpath,$ORIGIN/../lib", "-o", CLI])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
