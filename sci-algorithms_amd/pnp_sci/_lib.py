"""ctypes binding of the scigap C ABI (include/scigap.h).

The library is loaded lazily from this package directory.  There is no CPU
fallback: if ``libscigap.so`` is missing, or no HIP device is visible, every
entry point raises ``ScigapUnavailable``.
"""

from __future__ import annotations

import ctypes
import os
import threading

from .build import LIB_PATH

__all__ = ["ScigapError", "ScigapUnavailable", "Params", "KernelStat", "BatchItem",
           "BatchOptions", "Unrolled", "TuneResult", "lib", "check",
           "stream_ptr", "kernel_profile", "SYMBOLS", "METHOD_GAP", "METHOD_ADMM"]

METHOD_GAP = 0
METHOD_ADMM = 1

_ERRORS = {
    -1: "EINVAL", -2: "ENULL", -3: "EHIP", -4: "EWORKSPACE", -5: "EALIAS", -6: "EUNSUPPORTED",
}


class ScigapError(RuntimeError):
    """A scigap entry point returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__("scigap %s (%d): %s" % (_ERRORS.get(code, "?"), code, msg))
        self.code = code


class ScigapUnavailable(RuntimeError):
    """The HIP library or a HIP device is not available."""


class Params(ctypes.Structure):
    """Mirror of ``scigap_params``."""

    _fields_ = [
        ("H", ctypes.c_int64), ("W", ctypes.c_int64), ("T", ctypes.c_int64),
        ("method", ctypes.c_int), ("accelerate", ctypes.c_int), ("clip", ctypes.c_int),
        ("iters", ctypes.c_int),
        ("lambda_", ctypes.c_float), ("gamma", ctypes.c_float),
        ("tv_weight", ctypes.c_float), ("tv_eps", ctypes.c_float),
        ("tv_iter_max", ctypes.c_int),
        ("tv_weight_decay", ctypes.c_float), ("gamma_decay", ctypes.c_float),
        ("tile_rows", ctypes.c_int), ("engine", ctypes.c_int),
    ]


class BatchItem(ctypes.Structure):
    """Mirror of ``scigap_batch_item``."""

    _fields_ = [("H", ctypes.c_int64), ("W", ctypes.c_int64), ("T", ctypes.c_int64),
                ("y", ctypes.c_void_p), ("phi", ctypes.c_void_p), ("phi_sum", ctypes.c_void_p),
                ("x0", ctypes.c_void_p), ("x_out", ctypes.c_void_p)]


class BatchOptions(ctypes.Structure):
    """Mirror of ``scigap_batch_options``."""

    _fields_ = [("use_graph", ctypes.c_int), ("tile_rows", ctypes.c_int), ("grid", ctypes.c_int),
                ("order", ctypes.c_int)]


class Unrolled(ctypes.Structure):
    """Mirror of ``scigap_unrolled``."""

    _fields_ = [("H", ctypes.c_int64), ("W", ctypes.c_int64), ("T", ctypes.c_int64),
                ("batch", ctypes.c_int), ("iters", ctypes.c_int), ("tv_iter_max", ctypes.c_int),
                ("logits", ctypes.c_int), ("measure", ctypes.c_int)]


class TuneResult(ctypes.Structure):
    """Mirror of ``scigap_tune_result``."""

    _fields_ = [("tile_rows", ctypes.c_int), ("engine", ctypes.c_int), ("tv_depth", ctypes.c_int),
                ("candidates", ctypes.c_int), ("best_ms", ctypes.c_double),
                ("default_ms", ctypes.c_double)]


class KernelStat(ctypes.Structure):
    """Mirror of ``scigap_kernel_stat``."""

    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int),
                ("total_ms", ctypes.c_double)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int

# name -> (restype, argtypes); every function declared in include/scigap.h.
SYMBOLS = {
    "scigap_version": (_INT, []),
    "scigap_last_error": (ctypes.c_char_p, []),
    "scigap_forward": (_INT, [_P, _P, _P, _I64, _I64, _I64, _P]),
    "scigap_adjoint": (_INT, [_P, _P, _P, _I64, _I64, _I64, _P]),
    "scigap_phi_sum": (_INT, [_P, _P, _I64, _I64, _I64, _P]),
    "scigap_hwt_to_thw": (_INT, [_P, _P, _I64, _I64, _I64, _P]),
    "scigap_thw_to_hwt": (_INT, [_P, _P, _I64, _I64, _I64, _P]),
    "scigap_sq_error": (_INT, [_P, _P, _P, _I64, _I64, _I64, _INT, _P]),
    "scigap_ssim": (_INT, [_P, _P, _P, _I64, _I64, _I64, ctypes.c_float, _P]),
    "scigap_tv_workspace_bytes": (ctypes.c_size_t, [_I64, _I64, _I64, _INT]),
    "scigap_tv_chambolle": (_INT, [_P, _P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float,
                                   _INT, _P, _P, ctypes.c_size_t, _P]),
    "scigap_tv_workspace_bytes_depth": (ctypes.c_size_t, [_I64, _I64, _I64, _INT, _INT]),
    "scigap_tv_chambolle_depth": (_INT, [_P, _P, _I64, _I64, _I64, ctypes.c_float,
                                         ctypes.c_float, _INT, _INT, _P, _P, ctypes.c_size_t,
                                         _P]),
    "scigap_tv_noalloc_max_bytes": (ctypes.c_size_t, []),
    "scigap_tv_chambolle_noalloc": (_INT, [_P, _P, _I64, _I64, _I64, ctypes.c_float,
                                           ctypes.c_float, _INT, _P, _P]),
    "scigap_default_params": (None, [ctypes.POINTER(Params), _I64, _I64, _I64]),
    "scigap_solve_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Params)]),
    "scigap_solve": (_INT, [ctypes.POINTER(Params), _P, _P, _P, _P, _P, _P, _P, _P, _P,
                            ctypes.c_size_t, _P]),
    "scigap_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Params),
                                                       ctypes.POINTER(BatchItem), _INT,
                                                       ctypes.POINTER(BatchOptions)]),
    "scigap_batch_plan_create": (_INT, [ctypes.POINTER(_P), ctypes.POINTER(Params),
                                        ctypes.POINTER(BatchItem), _INT, _P, ctypes.c_size_t,
                                        ctypes.POINTER(BatchOptions), _P]),
    "scigap_batch_run": (_INT, [_P, _P]),
    "scigap_batch_plan_destroy": (_INT, [_P]),
    "scigap_unrolled_checkpoint_bytes": (ctypes.c_size_t, [ctypes.POINTER(Unrolled)]),
    "scigap_unrolled_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Unrolled)]),
    "scigap_unrolled_forward": (_INT, [ctypes.POINTER(Unrolled), _P, _P, _P, _P, _P, _P,
                                       ctypes.c_size_t, _P, ctypes.c_size_t, _P]),
    "scigap_unrolled_backward": (_INT, [ctypes.POINTER(Unrolled), _P, _P, _P, _P, _P,
                                        ctypes.c_size_t, _P, _P, _P, _P, _P, _P,
                                        ctypes.c_size_t, _P]),
    "scigap_autotune_solve": (_INT, [ctypes.POINTER(Params), _P, _P, _P, _P, _INT, _P,
                                     ctypes.c_size_t, _P, ctypes.POINTER(TuneResult)]),
    "scigap_autotune_tv": (_INT, [_P, _P, _I64, _I64, _I64, ctypes.c_float, ctypes.c_float, _INT,
                                  _P, ctypes.c_size_t, _P, ctypes.POINTER(TuneResult)]),
    "scigap_tune_clear": (_INT, []),
    "scigap_profile_begin": (_INT, []),
    "scigap_profile_end": (_INT, [ctypes.POINTER(KernelStat), _INT, ctypes.POINTER(_INT)]),
}


def kernel_profile(fn):
    """Run fn() with library kernel timing on; returns {kind: (launches, total_ms)}."""
    L = lib()
    check(L.scigap_profile_begin())
    try:
        fn()
    finally:
        buf = (KernelStat * 64)()
        n = _INT(0)
        check(L.scigap_profile_end(buf, 64, ctypes.byref(n)))
    return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(min(n.value, 64))}

_lock = threading.Lock()
_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libscigap.so and bind every symbol (no GPU needed to load)."""
    if not os.path.exists(path):
        raise ScigapUnavailable(
            "%s is not built; run __graft_entry__.build() or python -m pnp_sci.build" % path)
    dll = ctypes.CDLL(path)
    for name, (res, args) in SYMBOLS.items():
        fn = getattr(dll, name)
        fn.restype = res
        fn.argtypes = args
    return dll


def lib() -> ctypes.CDLL:
    """The loaded library, after checking that a HIP device is present."""
    global _lib
    with _lock:
        if _lib is None:
            import torch
            if not torch.cuda.is_available():
                raise ScigapUnavailable("no HIP device visible: the scigap GPU path cannot run")
            _lib = load_library()
        return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = _lib.scigap_last_error().decode(errors="replace") if _lib else ""
        raise ScigapError(rc, msg)


def stream_ptr(stream=None) -> int:
    """hipStream_t of a torch stream (the current stream by default)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
