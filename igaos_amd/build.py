"""Builds libigaos_pdhg.so and the `igaos` CLI for gfx950 with hipcc.

Every translation unit is compiled with ``--offload-arch=gfx950`` and with
packed FP32 instructions disabled; objects go to ``igaos_amd/_build`` and
the library lands in-tree next to this file, so the prebuilt ``.so``
travels with the repository snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libigaos_pdhg.so")
CLI = os.path.join(HERE, "igaos")

LIB_SOURCES = ["mps.cpp", "repair.cpp", "crossover.cpp", "engine.cpp", "capi.cpp",
               "pdhg_kernels.hip", "exact.hip", "batch.hip", "batch_host.cpp",
               "adjoint.hip", "unroll.hip", "noalloc.hip", "transpose.hip", "repair_dev.hip",
               "small.hip"]
# per-source extra flags: the exact mode's kernels must not fuse a*b + c
EXTRA = {"exact.hip": ["-ffp-contract=off"]}
CLI_SOURCES = ["cli.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-Wno-unused-command-line-argument",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC] + FLAGS + EXTRA.get(src, []) + lang + ["-c", path, "-o", obj], verbose)
    return obj


def build(verbose=False, jobs=None):
    """Compile every source for gfx950 and link the library and the CLI."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs + CLI_SOURCES))
    lib_objs = objs[:len(srcs)]
    cli_objs = objs[len(srcs):]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp] + lib_objs,
         verbose)
    os.replace(tmp, LIB)
    _run([HIPCC, f"--offload-arch={ARCH}", "-o", CLI] + cli_objs +
         ["-L", HERE, "-ligaos_pdhg", "-Wl,-rpath,$ORIGIN"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
