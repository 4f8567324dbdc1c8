"""Builds libctc.so (HIP kernels + C ABI) and the `ctc` command-line tool
in-tree for gfx950 with hipcc.  Usage: python -m cutc_amd.build [--jobs N]"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libctc.so")
CLI = os.path.join(HERE, "ctc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["host_coo.cpp", "host_ciss.cpp", "model.cpp", "ccd.cpp", "batch.cpp", "mixed.cpp", "kernels.hip", "kernels_ccd.hip", "kernels_mixed.hip", "grad_noalloc.hip", "quirks.cpp"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unknown-pragmas",
          # packed FP32 ops are disabled: they are avoided on this hardware pool
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", f"-I{os.path.join(HERE, '..', 'include')}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src):
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if not src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    _run(cmd)
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "ctc_cli.cpp")
    if os.path.exists(cli_src):
        tmp = CLI + ".tmp"
        _run([HIPCC, "-O2", "-std=c++17", f"-I{os.path.join(HERE, '..', 'include')}", cli_src, "-o", tmp,
              f"-L{HERE}", "-lctc", f"-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, CLI)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
