"""ctypes wrapper of the FP64 C oracle (gaptv_oracle.c) -- TEST INFRASTRUCTURE.

Arrays are taken in the reference's (H,W,T) layout and converted to the C
code's frame-major (T,H,W).  Build with ``make -C oracle``.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_gaptv.so")
if not os.path.exists(_LIB):
    raise ImportError("oracle/liboracle_gaptv.so is not built (make -C oracle)")
_dll = ctypes.CDLL(_LIB)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_dll.oracle_tv.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                           ctypes.c_double, ctypes.c_int, _I]
_dll.oracle_gap_tv.argtypes = [_D, _D, _D, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.c_double, ctypes.c_double, _D, _D, _I]

METHODS = {"gap": 0, "gap_plain": 1, "admm": 2}


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _thw(v):
    return np.ascontiguousarray(np.moveaxis(np.asarray(v, np.float64), -1, 0))


def tv(image, weight=0.1, eps=2e-4, n_iter_max=200):
    """Per-frame TV of an (H,W,T) volume; returns (out, returned iterations)."""
    src = _thw(image)
    T, H, W = src.shape
    out = np.empty_like(src)
    ret = np.zeros(T, dtype=np.int32)
    _dll.oracle_tv(_p(src), _p(out), H, W, T, weight, eps, n_iter_max,
                   ret.ctypes.data_as(_I))
    return np.moveaxis(out, 0, -1), ret


def gap_tv(y, mask, phi_sum, iters=50, method="gap", _lambda=1.0, gamma=0.01, clip=False,
           tv_weight=0.1, eps=2e-4, tv_iter_max=5, x0=None, truth=None, weight_decay=1.0,
           gamma_decay=1.0, exits=False):
    """FP64 GAP/ADMM-TV; returns x (H,W,T), plus sum of squared errors per
    iteration when `truth` is given, plus (last) the number of frames whose TV
    stopped early in each iteration when `exits`."""
    phi = _thw(mask)
    T, H, W = phi.shape
    yy = np.ascontiguousarray(y, np.float64)
    ps = np.ascontiguousarray(phi_sum, np.float64)
    xx0 = _thw(x0) if x0 is not None else None
    tr = _thw(truth) if truth is not None else None
    x = np.empty((T, H, W), dtype=np.float64)
    sq = np.zeros(iters, dtype=np.float64) if truth is not None else None
    ex = np.zeros(max(iters, 1), dtype=np.int32)
    _dll.oracle_gap_tv(_p(yy), _p(phi), _p(ps), _p(xx0), _p(x), H, W, T, METHODS[method],
                       int(clip), int(iters), float(_lambda), float(gamma), float(tv_weight),
                       float(eps), int(tv_iter_max), float(weight_decay), float(gamma_decay),
                       _p(tr), _p(sq), ex.ctypes.data_as(_I))
    xo = np.moveaxis(x, 0, -1)
    res = (xo,)
    if truth is not None:
        res += (sq,)
    if exits:
        res += (ex[:iters],)
    return res if len(res) > 1 else xo
