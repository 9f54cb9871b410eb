"""Build libgls.so (HIP, gfx950) in-tree with hipcc.

Usage: ``python -m cuvarbase_amd.build`` or ``cuvarbase_amd.build.build()``.
Every call recompiles all sources (no staleness checks) and writes
``cuvarbase_amd/libgls.so`` next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgls.so")
ARCH = os.environ.get("GLS_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "runtime.hip",
    "prep.hip",
    "nfft_passes.hip",
    "dirsum.hip",
    "misc.hip",
    "mh.hip",
    "plan.hip",
]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def _compile(src, obj, extra):
    cmd = [hipcc()] + HIPCC_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False, extra_flags=None, jobs=None):
    """Compile every source for gfx950 and link libgls.so (in-tree)."""
    extra = list(extra_flags or [])
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with tempfile.TemporaryDirectory(prefix="gls_build_") as tmp:
        objs = [os.path.join(tmp, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        with cf.ThreadPoolExecutor(jobs) as ex:
            futs = [ex.submit(_compile, s, o, extra) for s, o in zip(SOURCES, objs)]
            for f in futs:
                f.result()
        tmp_lib = os.path.join(tmp, "libgls.so")
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp_lib] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        shutil.move(tmp_lib, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
