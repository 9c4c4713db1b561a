"""Build libgraft_heff.so for gfx950 with hipcc (in-tree, next to this file).

Every source compiles to its own object in parallel (no device code crosses
a source file: each file launches only its own kernels), then one link."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["heff_runtime.hip", "zgemm_kernels.hip", "skinny_kernels.hip", "zgemm_lowp_kernels.hip",
           "vector_kernels.hip", "sector_kernels.hip", "zgemm_db_kernels.hip"]
OUT = os.path.join(HERE, "libgraft_heff.so")
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         # packed FP32 instructions are kept out of every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + "\n".join(msgs))


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    objs, cmds = [], []
    for s in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj])
    if verbose:
        print(f"{HIPCC} {' '.join(FLAGS)} -c <{len(SOURCES)} sources in parallel>", flush=True)
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as pool:
        list(pool.map(_run, cmds))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
