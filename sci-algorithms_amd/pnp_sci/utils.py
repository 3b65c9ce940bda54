"""Forward model and PSNR of PnP-SCI on the HIP device.

Mirrors ``PnP_SCI/python/utils.py:10-36`` (``A_``, ``At_``, ``psnr``): same
names, same ``(H,W,T)`` argument layout, results in the container kind of the
input (NumPy in, NumPy out; torch in, torch out).  The arithmetic runs in
fp32 in the scigap kernels.
"""

from __future__ import annotations

import math

from . import _lib
from .device import thw_to_hwt, to_device, to_device_thw, to_host_like, sq_error

__all__ = ["A_", "At_", "psnr", "make_operators", "SCIOperator"]


def _torch():
    import torch
    return torch


def A_(x, Phi):
    """y[h,w] = sum_t x[h,w,t]*Phi[h,w,t]  (utils.py:10-15) via scigap_forward."""
    torch = _torch()
    xd = to_device_thw(x, "x")
    pd = to_device_thw(Phi, "Phi")
    T, H, W = pd.shape
    if tuple(xd.shape) != (T, H, W):
        raise ValueError("x and Phi shapes differ: %s vs %s" % (tuple(x.shape), tuple(Phi.shape)))
    y = torch.empty((H, W), dtype=torch.float32, device=xd.device)
    L = _lib.lib()
    _lib.check(L.scigap_forward(xd.data_ptr(), pd.data_ptr(), y.data_ptr(), H, W, T,
                                _lib.stream_ptr()))
    return to_host_like(y, x)


def At_(y, Phi):
    """x[h,w,t] = y[h,w]*Phi[h,w,t]  (utils.py:17-26) via scigap_adjoint."""
    torch = _torch()
    yd = to_device(y, "y")
    pd = to_device_thw(Phi, "Phi")
    T, H, W = pd.shape
    if tuple(yd.shape) != (H, W):
        raise ValueError("y shape %s does not match Phi %s" % (tuple(y.shape), tuple(Phi.shape)))
    x = torch.empty((T, H, W), dtype=torch.float32, device=yd.device)
    L = _lib.lib()
    _lib.check(L.scigap_adjoint(yd.data_ptr(), pd.data_ptr(), x.data_ptr(), H, W, T,
                                _lib.stream_ptr()))
    return to_host_like(thw_to_hwt(x), y)


def psnr(ref, img):
    """PSNR with peak 1, 100 dB for an exact match (utils.py:28-36)."""
    a = to_device(ref, "ref")
    b = to_device(img, "img")
    if a.shape != b.shape:
        raise ValueError("shape mismatch %s vs %s" % (tuple(a.shape), tuple(b.shape)))
    a3 = a.reshape(1, 1, -1) if a.dim() != 3 else a
    b3 = b.reshape(1, 1, -1) if b.dim() != 3 else b
    s = float(sq_error(a3, b3).item())
    mse = s / a.numel()
    if mse == 0:
        return 100
    return 20 * math.log10(1.0 / math.sqrt(mse))


class SCIOperator:
    """Callable A or At bound to a mask held on the device as (T,H,W).

    ``gap_denoise``/``admm_denoise`` read the mask from ``scigap_phi`` instead
    of probing the callable, so building the operators with
    ``make_operators(mask)`` avoids the host round trip of the probe.
    """

    def __init__(self, mask, adjoint: bool):
        self.mask_hwt = mask
        self.scigap_phi = to_device_thw(mask, "mask")
        self.adjoint = adjoint

    def __call__(self, v):
        return At_(v, self.mask_hwt) if self.adjoint else A_(v, self.mask_hwt)


def make_operators(mask):
    """(A, At) for a (H,W,T) mask, like ``lambda x: A_(x, mask)``
    (pnp_sci_demo_kobe.py:74-75), carrying the device copy of the mask."""
    return SCIOperator(mask, adjoint=False), SCIOperator(mask, adjoint=True)
