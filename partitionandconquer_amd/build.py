"""Build the gfx950 library libpac.so (and the pac_main CLI) in-tree with hipcc.

    python -m partitionandconquer_amd.build          # incremental
    python -m partitionandconquer_amd.build --clean  # from scratch

Objects go to partitionandconquer_amd/_build/; the library and CLI next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpac.so")
CLI = os.path.join(HERE, "pac_main")
# test driver of the C++ facade (built only where the tests directory is present)
FACADE_CHECK_SRC = os.path.join(HERE, "..", "tests", "facade_check.cpp")
FACADE_CHECK = os.path.join(HERE, "..", "tests", "_bin", "facade_check")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PAC_OFFLOAD_ARCH", "gfx950")

LIB_SOURCES = ["pac_api.cpp", "pac_tune.cpp", "pac_prepare.hip", "pac_gen.hip",
               "pac_isect.hip", "pac_mlp.hip", "pac_multi.hip"]
CLI_SOURCES = ["pac_graph.cpp", "pac_main.cpp"]
HEADERS = ["pac_internal.hpp", "pac_graph.hpp", "pac_wait.hpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed-FP32 instructions are disabled for all device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
          "-I" + os.path.join(HERE, "..", "include")]


def _newer(src: str, dst: str, deps) -> bool:
    if not os.path.exists(dst):
        return True
    t = os.path.getmtime(dst)
    return any(os.path.getmtime(p) > t for p in [src, *deps] if os.path.exists(p))


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    deps = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(HERE, "..", "include", "pac.h")]
    if force or _newer(src, obj, deps):
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *COMMON, *lang, "-c", src, "-o", obj + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(obj + ".tmp", obj)
    return obj


def build(force: bool = False, jobs: int | None = None, cli: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    cli_srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
        cli_objs = list(ex.map(lambda s: _compile(s, force), cli_srcs)) if cli else []
    if force or any(_newer(o, LIB, []) for o in objs) or not os.path.exists(LIB):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", LIB + ".tmp", *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(LIB + ".tmp", LIB)
    if cli and len(cli_objs) == len(CLI_SOURCES):
        if force or any(_newer(o, CLI, [LIB]) for o in cli_objs) or not os.path.exists(CLI):
            cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", CLI + ".tmp", *cli_objs,
                   "-L" + HERE, "-lpac", "-Wl,-rpath,$ORIGIN"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
            os.replace(CLI + ".tmp", CLI)
    if cli and os.path.exists(FACADE_CHECK_SRC):
        graph_obj = os.path.join(BUILD, "pac_graph.cpp.o")
        if force or _newer(FACADE_CHECK_SRC, FACADE_CHECK,
                           [LIB, graph_obj, os.path.join(CSRC, "pac_graph.hpp")]):
            os.makedirs(os.path.dirname(FACADE_CHECK), exist_ok=True)
            obj = _compile(FACADE_CHECK_SRC, True)
            cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", FACADE_CHECK + ".tmp", obj, graph_obj,
                   "-L" + HERE, "-lpac", "-Wl,-rpath,$ORIGIN/../../partitionandconquer_amd"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
            os.replace(FACADE_CHECK + ".tmp", FACADE_CHECK)
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    if a.clean and os.path.isdir(BUILD):
        shutil.rmtree(BUILD)
    print(build(force=a.clean, jobs=a.j))
    return 0


if __name__ == "__main__":
    sys.exit(main())
