"""ctypes binding of the gfx950 HIP library (``libpytme_hip.so``).

The C-ABI is declared in ``include/pytme_hip.h``.  Every entry point returns
an int status; :func:`check` turns a non-zero status into ``ValueError``
(bad shape/argument), ``MemoryError`` (HBM allocation failure) or
``RuntimeError`` (any other HIP error).  There is no fallback: if the library
is missing, :func:`lib` raises ``ImportError``.
"""

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int64, c_void_p

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libpytme_hip.so")

TM_OK, TM_ERR_VALUE, TM_ERR_OOM, TM_ERR_HIP = 0, 1, 2, 3

c_int_p = POINTER(c_int)
c_int64_p = POINTER(c_int64)
c_double_p = POINTER(c_double)

# name -> (restype, argtypes); mirrors include/pytme_hip.h
SIGNATURES = {
    "tm_abi_version": (c_int, []),
    "tm_last_error": (c_char_p, []),
    "tm_device_count": (c_int, [c_int_p]),
    "tm_stream_sync": (c_int, [c_void_p]),
    "tm_timing_enable": (c_int, [c_int]),
    "tm_timing_names": (c_int, [c_char_p, c_int]),
    "tm_timing_query": (c_int, [c_char_p, c_double_p, c_int64_p, c_double_p]),
    # FFT
    "tm_fft_plan_create": (c_int, [c_int, c_int64_p, c_int64, POINTER(c_void_p)]),
    "tm_fft_plan_destroy": (c_int, [c_void_p]),
    "tm_fft_r2c": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_fft_c2r": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_fft_factor": (c_int, [c_int, c_int_p, c_int, c_int_p]),
    # rotation
    "tm_rotate_workspace": (c_int64, [c_int_p, c_int]),
    "tm_rotate_trilinear": (
        c_int,
        [
            c_void_p, c_void_p, c_int_p, c_void_p, c_void_p, c_double_p, c_double_p,
            c_int, c_void_p, c_void_p, c_int64_p, c_int64, c_void_p, c_void_p, c_void_p,
        ],
    ),
    # pointwise / reductions
    "tm_reduce_workspace": (c_int64, []),
    "tm_fill_f32": (c_int, [c_void_p, c_float, c_int64, c_void_p]),
    "tm_fill_i32": (c_int, [c_void_p, c_int, c_int64, c_void_p]),
    "tm_cmul_cc": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "tm_cmul_cr": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "tm_binary_f32": (
        c_int, [c_int, c_void_p, c_void_p, c_float, c_void_p, c_int64, c_int64, c_void_p]
    ),
    "tm_norm_scores": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_float, c_void_p, c_int64, c_int64, c_void_p],
    ),
    "tm_max_update": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int64, c_void_p]),
    "tm_sum_f64": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "tm_normalize_template": (
        c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_double, c_void_p, c_void_p]
    ),
    # fused batched scan
    "tm_scan_workspace_bytes": (c_int64, [c_int, c_int64_p, c_int_p, c_int, c_int]),
    "tm_scan_plan_create": (
        c_int,
        [
            c_int, c_int64_p, c_int_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
            c_int, c_double, c_int, POINTER(c_void_p),
        ],
    ),
    "tm_scan_plan_destroy": (c_int, [c_void_p]),
    "tm_scan_plan_set_target": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_scan_plan_use_graph": (c_int, [c_void_p, c_int]),
    "tm_scan_plan_graph_stats": (c_int, [c_void_p, c_int64_p, c_int64_p, c_int64_p]),
    "tm_scan_plan_set_tiles": (c_int, [c_void_p, c_int, c_int]),
    "tm_scan_plan_set_radix": (c_int, [c_void_p, c_int]),
    "tm_fft_radix_plan": (c_int, [c_int, c_int, c_int_p, c_int, c_int_p]),
    "tm_scan_group_create": (c_int, [POINTER(c_void_p), c_int, POINTER(c_void_p)]),
    "tm_scan_group_destroy": (c_int, [c_void_p]),
    "tm_scan_group_batch": (
        c_int, [c_void_p, POINTER(c_void_p), POINTER(c_void_p), c_int_p, POINTER(c_void_p), POINTER(c_void_p), c_void_p]
    ),
    # local FLC refinement
    "tm_flc_local_plan_create": (c_int, [c_void_p, c_void_p, c_int_p, c_int, POINTER(c_void_p)]),
    "tm_flc_local_plan_destroy": (c_int, [c_void_p]),
    "tm_flc_local_fwd": (c_int, [c_void_p, c_void_p, c_int64_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "tm_flc_local_fwd_jac": (c_int, [c_void_p, c_void_p, c_int64_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "tm_refine_prepare": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "tm_refine_adam": (
        c_int,
        [
            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
            c_int, c_double, c_double, c_double, c_double, c_double, c_int, c_double, c_double, c_void_p,
        ],
    ),
    "tm_flc_local_bwd": (
        c_int, [c_void_p, c_void_p, c_int64_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]
    ),
    "tm_refine_weights": (
        c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, c_void_p]
    ),
    "tm_refine_adam_grad": (
        c_int,
        [
            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, c_double, c_double,
            c_double, c_double, c_int, c_double, c_void_p,
        ],
    ),
    "tm_scan_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    # peak extraction
    "tm_peaks_max_filter": (c_int, [c_void_p, c_int64_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_peaks_write": (c_int, [c_void_p, c_void_p, c_int64_p, c_void_p, c_void_p]),
    "tm_topk_workspace_bytes": (c_int64, [c_int64, c_int64]),
    "tm_topk_f32": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_int64, c_void_p]),
    "tm_peak_filter_workspace_bytes": (c_int64, [c_int64]),
    "tm_peak_filter": (
        c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]
    ),
    "tm_peak_caller_fold_workspace_bytes": (c_int64, [c_int64, c_int64, c_int64]),
    "tm_peak_caller_fold_noalloc_max_bytes": (c_int64, []),
    "tm_peak_caller_fold_noalloc": (
        c_int,
        [
            c_void_p, c_int64, c_void_p, c_int64_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
            c_int64, c_int64, c_int64, c_int, c_float, c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p,
            c_void_p, c_void_p,
        ],
    ),
    "tm_peak_caller_fold": (
        c_int,
        [
            c_void_p, c_int64, c_void_p, c_int64_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
            c_int64, c_int64, c_int64, c_int, c_float, c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p,
            c_void_p, c_void_p, c_int64, c_void_p,
        ],
    ),
}

_LIB = None


def lib():
    """Load (once) and return the HIP library with typed entry points."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # torch's bundled HIP runtime must be the one the library binds to: load it
    # first so the library's libamdhip64 dependency resolves to that copy
    # instead of a second runtime instance.
    import torch  # noqa: F401

    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m pytme_amd._build` "
            "(there is no CPU fallback for the hip backend)"
        )
    handle = ctypes.CDLL(LIB_PATH)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _LIB = handle
    return _LIB


def check(status: int, what: str = "") -> None:
    if status == TM_OK:
        return
    msg = (lib().tm_last_error() or b"").decode(errors="replace")
    msg = f"{what}: {msg}" if what else msg
    if status == TM_ERR_VALUE:
        raise ValueError(msg)
    if status == TM_ERR_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


#: number of entry-point calls made through :func:`call`; a caller holding
#: a device copy of some array compares it to tell whether any library call
#: (which may have written that array) ran in between
CALLS = 0


def call(name: str, *args) -> None:
    global CALLS
    CALLS += 1
    check(getattr(lib(), name)(*args), name)


def i64_array(values):
    arr = (c_int64 * len(values))(*[int(v) for v in values])
    return arr


def i32_array(values):
    return (c_int * len(values))(*[int(v) for v in values])


def f64_array(values):
    return (c_double * len(values))(*[float(v) for v in values])


def timing_enable(on: bool) -> None:
    """Record HIP events around every library kernel launch (see tm_timing_*)."""
    call("tm_timing_enable", 1 if on else 0)


def timing_report() -> dict:
    """{kernel name: (total_ms, launches, algorithmic_bytes)} of recorded launches."""
    buf = ctypes.create_string_buffer(1 << 16)
    call("tm_timing_names", buf, len(buf))
    out = {}
    for name in buf.value.decode().split("\n"):
        if not name:
            continue
        ms, n, by = c_double(), c_int64(), c_double()
        call("tm_timing_query", name.encode(), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(by))
        out[name] = (ms.value, n.value, by.value)
    return out
