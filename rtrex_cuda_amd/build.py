"""Builds the native libraries of rtrex_cuda_amd in-tree.

    librtrex_hip.so   HIP kernels for gfx950 + the C ABI of include/rtrex.h
    librtrex_gen.so   seeded synthetic graph generators (host C++)
    librtrex_hostcheck.so  host builds of device routines, for CPU tests
    librtrex_bench.so calibration copy kernel for the HBM byte counters
    rtrex             the command-line driver (host C++ on the C ABI)

Run ``python -m rtrex_cuda_amd.build`` (or ``__graft_entry__.build()``).
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
HIP_LIB = os.path.join(PKG, "librtrex_hip.so")
GEN_LIB = os.path.join(PKG, "librtrex_gen.so")
CHECK_LIB = os.path.join(PKG, "librtrex_hostcheck.so")
BENCH_LIB = os.path.join(PKG, "librtrex_bench.so")
CLI = os.path.join(PKG, "rtrex")

HIP_SOURCES = ["api.hip", "sort.hip", "graph.hip", "count.hip", "clean.hip", "extract.hip",
               "serve.hip", "writers.hip", "memcache.hip", "ingest.hip"]

# IEEE FP64 throughout (no contraction into FMA, no fast-math): the supports
# must round exactly like the reference's separate multiply/add/divide.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(out: str = HIP_LIB) -> str:
    """Each source compiled to its own object in parallel (no device code is
    shared across sources: kernels are launched through their host stubs),
    then one link."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def compile_one(src: str) -> str:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([hipcc, *cflags, "-c", "-o", obj, os.path.join(CSRC, src)])
        return obj

    jobs = max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_bench(out: str = BENCH_LIB) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out + ".tmp"
    _run([hipcc, *HIPCC_FLAGS, "-o", tmp, os.path.join(CSRC, "bench_copy.hip")])
    os.replace(tmp, out)
    return out


def build_cli(out: str = CLI) -> str:
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", tmp, os.path.join(CSRC, "cli.cpp"),
          "-L", PKG, "-lrtrex_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_gen(out: str = GEN_LIB) -> str:
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", tmp,
          os.path.join(CSRC, "gen.cpp")])
    os.replace(tmp, out)
    return out


def build_hostcheck(out: str = CHECK_LIB) -> str:
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", "-pthread",
          "-o", tmp, os.path.join(CSRC, "hostcheck.cpp")])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_gen()
    build_hostcheck()
    build_hip()
    build_bench()
    build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "gen" in what:
        build_gen()
    elif "hip" in what:
        build_hip()
    elif "bench" in what:
        build_bench()
    elif "cli" in what:
        build_cli()
    else:
        build_all()
