"""Build libgls.so (HIP, gfx950) in-tree with hipcc.

Usage: ``python -m cuvarbase_amd.build`` or ``cuvarbase_amd.build.build()``.
Every call recompiles all sources (no staleness checks) and writes
``cuvarbase_amd/libgls.so`` next to this file.

The NFFT pass kernels are instantiated for the tiles listed in
``csrc/tiles_gen.hpp``, which every build regenerates from the measured tile
table ``tiles_gfx950.json`` (written by ``python -m cuvarbase_amd.autotune``)
plus the default rule of every size.  ``build(sweep=True)`` instead writes
``libgls_sweep.so`` with a kernel for every candidate tile, which the
autotuner times.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgls.so")
SWEEP_LIB = os.path.join(HERE, "libgls_sweep.so")
ARCH = os.environ.get("GLS_OFFLOAD_ARCH", "gfx950")
TILE_TABLE = os.path.join(HERE, f"tiles_{ARCH}.json")

SOURCES = [
    "runtime.hip",
    "prep.hip",
    "nfft_pass1.hip",
    "nfft_pass2.hip",
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


# ---- tile lists -------------------------------------------------------------
# (the same rules as csrc/tiles.hpp)
def pass1_sizes():
    """The pass-1 column lengths n1 of the grid-size catalogue (sizes.hpp)."""
    text = open(os.path.join(CSRC, "sizes.hpp")).read()
    body = text[text.index("#define GLS_PASS1_SIZES"):].split("\n\n")[0]
    return [int(v) for v in re.findall(r"X\((\d+)\)", body)]


ROW_LOG2 = range(6, 13)       # p2: rows of 64 .. 4096 cells
TILE_STEPS = (1, 2, 4, 8, 16)


def pass1_tile_ok(n1, c1):
    return c1 in TILE_STEPS and n1 * c1 <= 8192


def pass2_tile_ok(p2, cr):
    return cr in TILE_STEPS and 256 <= (cr << p2) <= 4096


def default_pass1_cols(n1):
    c = 1
    while c < 16 and n1 * c * 2 <= 4096:
        c *= 2
    return c


def default_pass2_rows(p2):
    return max(1, min(16, (1024 if p2 <= 8 else 2048) >> p2))


def radix_stages(length):
    """LDS radix stages of a length-L FFT (the radices of lds_fft.hpp, largest
    first)."""
    st, rem = 0, length
    while rem > 1:
        r = next(r for r in (16, 12, 9, 8, 4, 2, 5, 3) if rem % r == 0)
        rem //= r
        st += 1
    return st


def model_units(n1, p2):
    """The grid chooser's model cost per cell, in units of one pass over the
    cells (plan.hip model_cost before scaling)."""
    return 1.0 + 0.12 * (radix_stages(n1) + radix_stages(1 << p2))


def load_tile_table(path=TILE_TABLE):
    """[(n1, p2, c1, cr, ps)] of the measured table (empty if there is none):
    tiles and the measured picoseconds per grid cell of the two passes."""
    if not os.path.exists(path):
        return []
    rows = []
    for b in json.load(open(path))["buckets"]:
        n1, p2, c1, cr = int(b["n1"]), int(b["p2"]), int(b["c1"]), int(b["cr"])
        if not (pass1_tile_ok(n1, c1) and pass2_tile_ok(p2, cr)):
            raise ValueError(f"{path}: tile ({c1}, {cr}) of bucket ({n1}, {p2}) "
                             "is not a candidate")
        rows.append((n1, p2, c1, cr, round(float(b["ps_per_cell"]), 3)))
    return rows


def model_scale(table):
    """Picoseconds per model unit for buckets the table does not list: the
    median of measured cost / model units over the table (1 without one)."""
    if not table:
        return 1.0
    r = sorted(ps / model_units(n1, p2) for n1, p2, _, _, ps in table)
    return round(r[len(r) // 2], 4)


def tile_header(sweep=False, table=None):
    """Text of the generated tile header: the (n1, c1) and (p2, cr) pairs that
    get kernels, the measured table (empty for the sweep build) and the model
    scale."""
    sizes = pass1_sizes()
    table = [] if sweep else (load_tile_table() if table is None else table)
    if sweep:
        p1 = [(n1, c1) for n1 in sizes for c1 in TILE_STEPS if pass1_tile_ok(n1, c1)]
        p2 = [(p, cr) for p in ROW_LOG2 for cr in TILE_STEPS if pass2_tile_ok(p, cr)]
    else:
        p1 = sorted({(n1, default_pass1_cols(n1)) for n1 in sizes}
                    | {(t[0], t[2]) for t in table})
        p2 = sorted({(p, default_pass2_rows(p)) for p in ROW_LOG2}
                    | {(t[1], t[3]) for t in table})

    def macro(name, items, per_line=6):
        lines = [" ".join(f"X({', '.join(map(str, t))})" for t in items[i:i + per_line])
                 for i in range(0, len(items), per_line)]
        body = " \\\n    ".join(lines)
        return f"#define {name}(X) \\\n    {body}\n" if items else f"#define {name}(X)\n"

    src = "the candidate tiles (tile sweep build)" if sweep else \
        f"{os.path.basename(TILE_TABLE)} and the default rules"
    return ("// Generated by cuvarbase_amd/build.py from " + src + "; do not edit.\n"
            "#pragma once\n\n"
            "// (n1, c1): pass-1 kernels of this library\n" + macro("GLS_PASS1_TILES", p1) +
            "\n// (p2, cr): pass-2 kernels of this library\n" + macro("GLS_PASS2_TILES", p2) +
            "\n// measured per size bucket: (n1, p2, c1, cr, picoseconds per grid cell)\n" +
            macro("GLS_TILE_TABLE", sorted(table), per_line=3) +
            "\n// picoseconds per model unit for buckets the table does not list\n"
            f"#define GLS_MODEL_PS_PER_UNIT {model_scale(table)!r}\n")


def write_tile_header(path, sweep=False):
    text = tile_header(sweep=sweep)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def build(verbose=False, extra_flags=None, jobs=None, sweep=False):
    """Compile every source for gfx950 and link libgls.so (in-tree); with
    sweep=True, libgls_sweep.so with a kernel for every candidate tile."""
    extra = list(extra_flags or [])
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    out_lib = SWEEP_LIB if sweep else LIB
    with tempfile.TemporaryDirectory(prefix="gls_build_") as tmp:
        if sweep:
            write_tile_header(os.path.join(tmp, "tiles_sweep.hpp"), sweep=True)
            extra += ["-DGLS_TILE_SWEEP", "-I", tmp]
        else:
            write_tile_header(os.path.join(CSRC, "tiles_gen.hpp"))
        objs = [os.path.join(tmp, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        with cf.ThreadPoolExecutor(jobs) as ex:
            futs = [ex.submit(_compile, s, o, extra) for s, o in zip(SOURCES, objs)]
            for f in futs:
                f.result()
        tmp_lib = os.path.join(tmp, os.path.basename(out_lib))
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp_lib] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        # copy next to the target, then rename over it (a reader never sees a
        # partly written library)
        staged = out_lib + ".tmp"
        shutil.copyfile(tmp_lib, staged)
        os.replace(staged, out_lib)
    if verbose:
        print(f"built {out_lib}")
    return out_lib


if __name__ == "__main__":
    build(verbose=True, sweep="--sweep" in sys.argv[1:])
    sys.exit(0)
