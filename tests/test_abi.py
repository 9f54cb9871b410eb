"""The C ABI: libgls.so loads and exports every function include/gls.h declares,
and the ctypes structure layouts match the C compiler's (no device calls)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gls.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gls_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gls_lomb_execute" in names and "gls_nfft_execute" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol(gls_lib):
    from cuvarbase_amd import _lib

    for name in declared_functions():
        assert hasattr(gls_lib, name), f"{name} declared in gls.h but not exported"
        assert name in _lib.EXPORTED_SYMBOLS, f"{name} has no ctypes prototype"


def test_ctypes_struct_layout_matches_c():
    from cuvarbase_amd import _lib

    fields = {
        "gls_curve": [f[0] for f in _lib.gls_curve._fields_],
        "gls_params": [f[0] for f in _lib.gls_params._fields_],
        "gls_nfft_item": [f[0] for f in _lib.gls_nfft_item._fields_],
        "gls_mh_curve": [f[0] for f in _lib.gls_mh_curve._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
             "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.check_call(["gcc", "-std=c99", src, "-o", exe])
        out = subprocess.check_output([exe], text=True)
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for st in fields:
        cls = getattr(_lib, st)
        assert int(got[f"{st} size"]) == C.sizeof(cls), st
        for f in fields[st]:
            assert int(got[f"{st}.{f}"]) == getattr(cls, f).offset, f"{st}.{f}"


def test_status_codes_match_header():
    from cuvarbase_amd import _lib

    txt = open(HEADER).read()
    for name in ("OK", "E_INVALID", "E_HIP", "E_OOM", "E_NOT_INIT", "E_UNSUPPORTED", "E_DATA"):
        m = re.search(rf"#define GLS_{name}\s+(-?\d+)", txt)
        assert m, name
        assert int(m.group(1)) == getattr(_lib, f"GLS_{name}")
    for name, val in (("MODE_STANDARD", 0), ("MODE_FLOATING_MEAN", 1), ("MODE_WINDOW", 2),
                      ("INPUT_RAW", 0), ("INPUT_WEIGHTED", 1), ("METHOD_NFFT", 0),
                      ("METHOD_DIRECT", 1)):
        m = re.search(rf"#define GLS_{name}\s+(-?\d+)", txt)
        assert int(m.group(1)) == val == getattr(_lib, name)


def test_static_scratch_size_matches_header():
    from cuvarbase_amd import _lib

    m = re.search(r"#define GLS_STATIC_SCRATCH_BYTES \((\d+)u << (\d+)\)", open(HEADER).read())
    assert m
    assert int(m.group(1)) << int(m.group(2)) == _lib.STATIC_SCRATCH_BYTES


def test_missing_library_fails_loudly(tmp_path):
    from cuvarbase_amd import _lib

    with pytest.raises(_lib.GLSLibraryError):
        saved = _lib._lib
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "nope.so"))
        finally:
            _lib._lib = saved
