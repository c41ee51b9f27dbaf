"""Builds libsmccat_hip.so (HIP kernels + C-ABI, gfx950) and the C++ facade test.

Run as ``python smc-cat_amd/smccat/build.py``; __graft_entry__.build() calls build_all().
Every build recompiles from source (seconds per file) into smc-cat_amd/_build/ and
links the library in-tree, next to this file, so it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT_PKG = os.path.dirname(PKG)                 # smc-cat_amd/
REPO = os.path.dirname(ROOT_PKG)
CSRC = os.path.join(ROOT_PKG, "csrc")
BUILD = os.path.join(ROOT_PKG, "_build")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libsmccat_hip.so")
FACADE_EXE = os.path.join(PKG, "facade_selftest")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 VALU instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-I", INCLUDE, "-I", CSRC,
          "-Wall", "-Wno-unused-result"]

SOURCES = ["engine.cpp", "gamma.cpp", "lg_table.cpp", "k_eigen.hip", "k_prune.hip", "k_prune_mfma.hip",
           "k_smc.hip", "k_site_class.hip", "gibbs.cpp", "mcmc.cpp", "sampler.cpp", "k_mlp.hip", "mlp.cpp", "noalloc.cpp", "k_splits.hip", "k_tune.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_facade_test(verbose=False):
    """C++ facade (reference Particle API) + its self-test executable."""
    os.makedirs(BUILD, exist_ok=True)
    exe = FACADE_EXE
    cmd = ["g++", "-O2", "-std=c++17", "-I", INCLUDE, "-I", CSRC, "-I", os.path.join(CSRC, "host"),
           os.path.join(CSRC, "host", "particle.cpp"), os.path.join(CSRC, "host", "facade_selftest.cpp"),
           "-L", PKG, "-lsmccat_hip", "-Wl,-rpath,$ORIGIN", "-o", exe]
    _run(cmd)
    if verbose:
        print("built", exe)
    return exe


def build_oracle(verbose=False):
    out = _run(["make", "-s", "-C", os.path.join(REPO, "oracle")])
    if verbose and out.strip():
        print(out)
    return os.path.join(REPO, "oracle", "libsmccat_oracle.so")


def build_all(verbose=True):
    build_library(verbose)
    if os.path.exists(os.path.join(CSRC, "host", "facade_selftest.cpp")):
        build_facade_test(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
