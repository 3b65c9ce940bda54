"""Device containers for the scigap C ABI: PyTorch-ROCm tensors hold the
memory and supply the stream; every computation is a scigap call.

Reference arrays are NumPy ``(H,W,T)`` (PnP_SCI/python/utils.py:10-26); the
library works on frame-major ``(T,H,W)``.  The conversions here run the
library's own transpose kernels on the device.
"""

from __future__ import annotations

import numpy as np

from . import _lib

__all__ = ["is_torch", "to_device", "to_host_like", "hwt_to_thw", "thw_to_hwt", "sq_error",
           "DEVICE"]

DEVICE = "cuda"


def _torch():
    import torch
    return torch


def is_torch(a) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is part of the image
        return False
    return isinstance(a, torch.Tensor)


def to_device(a, name: str = "array"):
    """Contiguous float32 copy (or view) of `a` on the HIP device."""
    torch = _torch()
    if is_torch(a):
        t = a.detach()
        if t.device.type != "cuda":
            t = t.to(DEVICE)
        return t.to(torch.float32).contiguous()
    arr = np.asarray(a)
    if arr.dtype.kind not in "fiub":
        raise TypeError("%s must be numeric, got %s" % (name, arr.dtype))
    return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).to(DEVICE)


def to_host_like(t, like):
    """Return device tensor `t` in the container kind and dtype of `like`."""
    torch = _torch()
    if is_torch(like):
        dt = like.dtype if like.dtype.is_floating_point else torch.float32
        return t.to(device=like.device, dtype=dt)
    dt = np.asarray(like).dtype
    if dt.kind != "f":
        dt = np.float32
    return t.cpu().numpy().astype(dt, copy=False)


def hwt_to_thw(v):
    """(H,W,T) device tensor -> new (T,H,W) tensor (scigap_hwt_to_thw)."""
    torch = _torch()
    H, W, T = v.shape
    out = torch.empty((T, H, W), dtype=torch.float32, device=v.device)
    L = _lib.lib()
    _lib.check(L.scigap_hwt_to_thw(v.data_ptr(), out.data_ptr(), H, W, T, _lib.stream_ptr()))
    return out


def thw_to_hwt(v):
    """(T,H,W) device tensor -> new (H,W,T) tensor (scigap_thw_to_hwt)."""
    torch = _torch()
    T, H, W = v.shape
    out = torch.empty((H, W, T), dtype=torch.float32, device=v.device)
    L = _lib.lib()
    _lib.check(L.scigap_thw_to_hwt(v.data_ptr(), out.data_ptr(), H, W, T, _lib.stream_ptr()))
    return out


def sq_error(a, b, per_frame: bool = False):
    """sum((a-b)^2) in fp64 over (T,H,W) tensors (per frame if requested)."""
    torch = _torch()
    T, H, W = a.shape
    out = torch.empty(T if per_frame else 1, dtype=torch.float64, device=a.device)
    L = _lib.lib()
    _lib.check(L.scigap_sq_error(a.data_ptr(), b.data_ptr(), out.data_ptr(), H, W, T,
                                 1 if per_frame else 0, _lib.stream_ptr()))
    return out
