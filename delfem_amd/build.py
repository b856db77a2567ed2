"""Builds libdelfem_amd.so (HIP, gfx950) in-tree with hipcc, and the C++
program that drives the drop-in classes of include/delfem_amd/ (tests/cpp/)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdelfem_amd.so")

SOURCES = ["dfm_host.cpp", "dfm_api.hip", "dfm_solver.hip", "dfm_batch.hip", "dfm_complex.hip", "dfm_bmat.hip", "dfm_mseg.hip"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


SHIM_PROGRAMS = ["shim_test", "zshim_test", "msshim_test"]  # tests/cpp/<name>.cpp -> tests/cpp/<name>


def build_shim_test(verbose=False):
    """The C++ programs of the drop-in tests (real and complex classes),
    linked against the library."""
    cxx = os.environ.get("CXX", "g++")
    out = []
    for name in SHIM_PROGRAMS:
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        binp = os.path.join(ROOT, "tests", "cpp", name)
        tmp = binp + ".tmp"
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", tmp,
               "-L", HERE, "-ldelfem_amd", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../delfem_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, binp)
        out.append(binp)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_shim_test(verbose="-v" in sys.argv)
    print(LIB)
