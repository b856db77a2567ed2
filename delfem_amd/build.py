"""Builds libdelfem_amd.so (HIP, gfx950) in-tree with hipcc, and the C++
program that drives the drop-in classes of include/delfem_amd/ (tests/cpp/)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdelfem_amd.so")

SOURCES = ["dfm_host.cpp", "dfm_api.hip", "dfm_solver.hip", "dfm_batch.hip", "dfm_complex.hip", "dfm_bmat.hip", "dfm_mseg.hip", "dfm_flex.hip"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


OBJ_DIR = os.path.join(ROOT, "_build")


def build(verbose=False, jobs=None):
    """Compiles every source to an object (in parallel, `jobs` at a time) and
    links the shared library; always recompiles."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]

    def compile_one(src):
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *cflags, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = jobs or min(len(SOURCES), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


SHIM_PROGRAMS = ["shim_test", "zshim_test", "msshim_test", "poisson_dropin", "flex_dropin"]  # tests/cpp/<name>.cpp -> tests/cpp/<name>


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
