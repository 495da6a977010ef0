"""Build libdakc_amd.so (HIP, gfx950), the dakc CLI and the CPU oracle in-tree.

    python -m dakc_amd.build          # or __graft_entry__.build()

Outputs:
    dakc_amd/libdakc_amd.so        the product: C ABI of include/dakc.h
    dakc_amd/dakc                  CLI driver (file in, k, table out)
    dakc_amd/hpp_example           smallest driver of include/dakc.hpp
    oracle/_build/libdakc_oracle.so  CPU oracle (test infrastructure)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libdakc_amd.so"
CLI = PKG / "dakc"
HPP_EXAMPLE = PKG / "hpp_example"
ORACLE_LIB = ROOT / "oracle" / "_build" / "libdakc_oracle.so"

HIP_SOURCES = ["kcount.hip", "api.hip", "serve.hip", "batch.hip", "post.hip", "input.hip", "noalloc.hip", "phases.hip", "hitter.hip"]
ARCH = os.environ.get("DAKC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions stay disabled in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (rc={r.returncode})")


def build_lib(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: str) -> Path:
        obj = BUILD / (src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], verbose)
        return obj

    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cli(verbose: bool = False) -> Path:
    for src, dst in (("dakc_main.cpp", CLI), ("hpp_example.cpp", HPP_EXAMPLE)):
        tmp = dst.with_name(dst.name + ".tmp")
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(PKG / "cli" / src),
              "-o", str(tmp), f"-L{PKG}", "-ldakc_amd", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(tmp, dst)
    return CLI


def build_oracle(verbose: bool = False) -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")], verbose)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_lib(verbose)
    build_cli(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(f"built {LIB}, {CLI}, {ORACLE_LIB}")
