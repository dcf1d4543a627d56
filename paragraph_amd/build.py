"""Compiles the HIP sources into paragraph_amd/libparagraph_hip.so (gfx950)
and the host stages into paragraph_amd/libparagraph_host.so.

Run as `python -m paragraph_amd.build`.  hipcc cross-compiles without a GPU,
so this works on a CPU-only machine; the shared library then travels with the
source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libparagraph_hip.so")
ARCH = os.environ.get("PG_OFFLOAD_ARCH", "gfx950")

# pg_search_*.hip hold the search kernels, one evaluator family each, so
# that the largest instantiation sets compile in parallel
HIP_SOURCES = ["pg_search_f32_16.hip", "pg_search_f32_8.hip",
               "pg_search_f32_4.hip", "pg_search_bf16.hip",
               "pg_search_fp8.hip", "pg_kernels.hip", "pg_search_f32_32.hip",
               "pg_topk.hip", "pg_capi.cpp"]
# parallel compiles (each hipcc holds a few GB at most)
JOBS = max(1, min(8, os.cpu_count() or 1))

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    hipcc = _hipcc()
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, "-x", "hip", f"--offload-arch={ARCH}", *COMMON,
               "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        return src, obj, subprocess.run(cmd, capture_output=True, text=True)

    with ThreadPoolExecutor(max_workers=JOBS) as pool:
        results = list(pool.map(compile_one, HIP_SOURCES))
    objs = []
    for src, obj, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST_LIB = os.path.join(HERE, "libparagraph_host.so")
HOST_SOURCES = ["pg_host_stage.cpp"]


def build_host(verbose=False):
    """The host stages (entry point, supply / bipartite graphs, final lists,
    index file) into libparagraph_host.so: plain C++ with OpenMP, explicit
    FMAs (no contraction)."""
    cxx = shutil.which("g++") or "g++"
    tmp = HOST_LIB + ".tmp"
    cmd = [cxx, "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
           "-ffp-contract=off", "-mavx2", "-mfma", "-Wall", "-Wextra",
           *[os.path.join(CSRC, s) for s in HOST_SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("host stage build failed")
    os.replace(tmp, HOST_LIB)
    build_tool(verbose)
    return HOST_LIB


TOOL = os.path.join(HERE, "pg_build_index")
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")


def build_tool(verbose=False):
    """pg_build_index, the command-line whole build on the C++ front end
    (include/paragraph/Builder.hpp), linked against both libraries."""
    cxx = shutil.which("g++") or "g++"
    tmp = TOOL + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-pthread",
           os.path.join(CSRC, "pg_build_index.cpp"), "-o", tmp,
           "-L" + HERE, "-lparagraph_hip", "-lparagraph_host",
           "-L" + ROCM_LIB, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + ROCM_LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("pg_build_index build failed")
    os.replace(tmp, TOOL)
    return TOOL


if __name__ == "__main__":
    print(build(verbose=True))
