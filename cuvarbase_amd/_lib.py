"""ctypes binding of libgls.so (the HIP C ABI declared in include/gls.h).

The library is loaded from this package directory only.  There is no fallback:
if ``libgls.so`` is missing or cannot be loaded, every entry point raises
``GLSLibraryError``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; CUVARBASE_AMD_LIBGLS names another build of the same ABI
LIB_PATH = os.environ.get("CUVARBASE_AMD_LIBGLS") or os.path.join(HERE, "libgls.so")

GLS_OK = 0
GLS_E_INVALID = -1
GLS_E_HIP = -2
GLS_E_OOM = -3
GLS_E_NOT_INIT = -4
GLS_E_UNSUPPORTED = -5
GLS_E_DATA = -6

MODE_STANDARD = 0
MODE_FLOATING_MEAN = 1
MODE_WINDOW = 2

INPUT_RAW = 0
INPUT_WEIGHTED = 1

METHOD_NFFT = 0
METHOD_DIRECT = 1
METHOD_DIRECT_CUSTOM = 2
WINDOW_GAUSSIAN = 0
WINDOW_ES = 1

PRECISION_FP64_COORDS = 0
PRECISION_FP32 = 1
SPREAD_FLOAT = 0
SPREAD_FIXED = 1
SPREAD_AUTO = 2

STATIC_SCRATCH_BYTES = 128 << 20   # GLS_STATIC_SCRATCH_BYTES


class GLSError(RuntimeError):
    """A libgls call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libgls error {code}: {msg}")
        self.code = code


class GLSOutOfMemory(GLSError, MemoryError):
    pass


class GLSLibraryError(ImportError):
    pass


class gls_curve(C.Structure):
    _fields_ = [("obs_offset", C.c_int64), ("n_obs", C.c_int32), ("nf", C.c_int32),
                ("f0", C.c_double), ("df", C.c_double), ("out_offset", C.c_int64),
                ("ybar", C.c_double), ("yy", C.c_double)]


class gls_params(C.Structure):
    _fields_ = [("sigma", C.c_double), ("m", C.c_int32), ("mode", C.c_int32),
                ("reg_c", C.c_double), ("reg_s", C.c_double),
                ("input_kind", C.c_int32), ("method", C.c_int32),
                ("window", C.c_int32), ("window_beta", C.c_float),
                ("grid_chunk_bytes", C.c_int64), ("point_capacity", C.c_int64),
                ("precision", C.c_int32), ("spread", C.c_int32)]


class gls_mh_curve(C.Structure):
    _fields_ = [("sums_offset", C.c_int64), ("out_offset", C.c_int64), ("nf", C.c_int32),
                ("pad", C.c_int32), ("yy", C.c_double), ("ybar", C.c_double),
                ("sum_w", C.c_double), ("lam", C.c_double)]


class gls_nfft_item(C.Structure):
    _fields_ = [("obs_offset", C.c_int64), ("n_obs", C.c_int32), ("nf", C.c_int32),
                ("f0", C.c_double), ("df", C.c_double), ("out_offset", C.c_int64)]


_vp = C.c_void_p
_PROTOS = {
    "gls_init": [C.c_int],
    "gls_shutdown": [],
    "gls_device_count": [C.POINTER(C.c_int)],
    "gls_device_info": [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_size_t)],
    "gls_version": [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "gls_stream_create": [C.POINTER(_vp)],
    "gls_stream_destroy": [_vp],
    "gls_stream_synchronize": [_vp],
    "gls_stream_query": [_vp, C.POINTER(C.c_int)],
    "gls_event_create": [C.POINTER(_vp)],
    "gls_event_destroy": [_vp],
    "gls_event_record": [_vp, _vp],
    "gls_event_synchronize": [_vp],
    "gls_event_query": [_vp, C.POINTER(C.c_int)],
    "gls_event_elapsed_ms": [_vp, _vp, C.POINTER(C.c_float)],
    "gls_device_synchronize": [],
    "gls_malloc": [C.POINTER(_vp), C.c_size_t],
    "gls_free": [_vp],
    "gls_host_alloc": [C.POINTER(_vp), C.c_size_t],
    "gls_host_free": [_vp],
    "gls_memcpy_h2d_async": [_vp, _vp, C.c_size_t, _vp],
    "gls_memcpy_d2h_async": [_vp, _vp, C.c_size_t, _vp],
    "gls_memcpy_d2d_async": [_vp, _vp, C.c_size_t, _vp],
    "gls_memcpy_kernel_async": [_vp, _vp, C.c_size_t, _vp],
    "gls_memset_async": [_vp, C.c_int, C.c_size_t, _vp],
    "gls_mem_info": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "gls_lomb_plan_create": [C.POINTER(_vp), C.POINTER(gls_curve), C.c_int32,
                             C.POINTER(gls_params)],
    "gls_plan_destroy": [_vp],
    "gls_plan_workspace_bytes": [_vp, C.POINTER(C.c_size_t)],
    "gls_plan_output_count": [_vp, C.POINTER(C.c_int64)],
    "gls_plan_grid_size": [_vp, C.c_int32, C.POINTER(C.c_int64)],
    "gls_plan_curve_geometry": [_vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32), C.POINTER(C.c_double)],
    "gls_plan_curve_tiles": [_vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "gls_plan_curve_spread": [_vp, C.c_int32, C.POINTER(C.c_int32)],
    "gls_launch_items_set": [C.c_int64],
    "gls_plan_stage_bucket": [_vp, C.c_int32] + [C.POINTER(C.c_int32)] * 6,
    "gls_tile_set": [C.c_int32, C.c_int32, C.c_int32, C.c_int32],
    "gls_tile_get": [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "gls_grid_cost_set": [C.c_int32, C.c_int32, C.c_double],
    "gls_grid_cost_get": [C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)],
    "gls_lomb_execute": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "gls_lomb_execute_f64": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "gls_lomb_execute_static": [_vp, _vp, _vp, _vp, _vp, _vp],
    "gls_lomb_execute_custom": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, _vp, _vp],
    "gls_mh_power": [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_int32, _vp],
    "gls_plan_set_fork": [_vp, C.c_int32],
    "gls_plan_graph_create": [_vp, _vp, _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp],
    "gls_graph_launch": [_vp, _vp],
    "gls_graph_destroy": [_vp],
    "gls_plan_rebind": [_vp, C.POINTER(gls_curve), C.c_int32],
    "gls_graph_update": [_vp, _vp, _vp],
    "gls_nfft_plan_create": [C.POINTER(_vp), C.POINTER(gls_nfft_item), C.c_int32,
                             C.c_double, C.c_int32],
    "gls_nfft_execute": [_vp, _vp, _vp, C.c_int32, _vp, _vp, _vp],
    "gls_gaussian_grid": [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                          C.c_double, C.c_double, C.c_double, _vp, _vp],
    "gls_topk": [_vp, _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp],
    "gls_plan_enable_timing": [_vp, C.c_int32],
    "gls_plan_stage_count": [_vp, C.POINTER(C.c_int32)],
    "gls_plan_stage_times": [_vp, _vp, _vp, _vp, _vp, C.c_int32],
}

EXPORTED_SYMBOLS = tuple(_PROTOS) + ("gls_last_error",)

_lib = None
_lock = threading.Lock()
_initialised_device = None


def load(path=None):
    """Load libgls.so (once).  Raises GLSLibraryError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise GLSLibraryError(
                f"{p} not found: build it with `python -m cuvarbase_amd.build` "
                "(there is no CPU fallback)")
        try:
            lib = C.CDLL(p)
        except OSError as e:  # pragma: no cover - depends on the machine
            raise GLSLibraryError(f"cannot load {p}: {e}") from e
        lib.gls_last_error.restype = C.c_char_p
        lib.gls_last_error.argtypes = []
        for name, args in _PROTOS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = lib
        return lib


def check(rc):
    if rc != GLS_OK:
        msg = load().gls_last_error().decode(errors="replace")
        if rc == GLS_E_OOM:
            raise GLSOutOfMemory(rc, msg)
        raise GLSError(rc, msg)
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))


def init(device=0):
    """Create the device context lazily (replaces import-time context creation
    of the reference, cuvarbase/__init__.py:2)."""
    global _initialised_device
    if _initialised_device == device:
        return
    call("gls_init", int(device))
    _initialised_device = device


def device_count():
    n = C.c_int(0)
    rc = load().gls_device_count(C.byref(n))
    return n.value if rc == GLS_OK else 0
