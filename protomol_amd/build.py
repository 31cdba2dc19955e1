"""Build libprotomol_amd.so (the HIP kernels + C ABI) in-tree for gfx950.

Usage: python -m protomol_amd.build  (or build() from __graft_entry__)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libprotomol_amd.so")
# the pair-kernel sources first: they take longest
SOURCES = ["pm_pair_lj_ewald.hip", "pm_pair_lj_coulomb.hip", "pm_pair_lj.hip", "pm_pair_coulomb.hip",
           "pm_pair_ewald.hip", "pm_kernels.hip", "pm_local.hip", "pm_gb.hip", "pm_ewald.hip", "pm_api.cpp"]
# per-source flags: the local terms are built without machine-level LICM
# (their FP64 library constants would otherwise stay in registers across loops)
EXTRA = {"pm_local.hip": ["-mllvm", "-disable-machine-licm"]}
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")

# packed-FP32 instructions are kept out of the generated code on purpose
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 0) -> str:
    jobs = jobs or max(1, min(len(SOURCES), os.cpu_count() or 1))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmd = [cc, *lang, *HIPFLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            time.sleep(0.1)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    t = time.time()
    print(build(verbose="-v" in sys.argv), f"{time.time() - t:.1f}s")
