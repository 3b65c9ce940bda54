"""Device containers for the scigap C ABI: PyTorch-ROCm tensors hold the
memory and supply the stream; every computation is a scigap call.

Reference arrays are NumPy ``(H,W,T)`` (PnP_SCI/python/utils.py:10-26); the
library works on frame-major ``(T,H,W)``.  The conversions here run the
library's own transpose kernels on the device.
"""

from __future__ import annotations

import threading

import numpy as np

from . import _lib

__all__ = ["is_torch", "to_device", "to_device_thw", "to_host_like", "hwt_to_thw", "thw_to_hwt",
           "sq_error", "DEVICE"]

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


# NumPy arrays from this size on cross PCIe through two pinned staging buffers
# (allocated once, 16 MiB each): the host copy into one buffer, including the
# cast to float32, overlaps the DMA from the other.  Pageable copies of large
# arrays run several times slower.
_STAGE_BYTES = 16 << 20
_STAGE_MIN_BYTES = 4 << 20
_stage_lock = threading.Lock()
_stage_bufs = None


def _staging():
    global _stage_bufs
    if _stage_bufs is None:
        torch = _torch()
        _stage_bufs = [torch.empty(_STAGE_BYTES // 4, dtype=torch.float32, pin_memory=True)
                       for _ in range(2)]
    return _stage_bufs


def _h2d_staged(arr: np.ndarray):
    """float32 device copy of a (large) NumPy array through pinned staging."""
    torch = _torch()
    out = torch.empty(arr.shape, dtype=torch.float32, device=DEVICE)
    src = np.ravel(arr, order="C")
    dst = out.view(-1)
    n, chunk = src.size, _STAGE_BYTES // 4
    with _stage_lock:
        bufs = _staging()
        done = [None, None]
        for i, off in enumerate(range(0, n, chunk)):
            k = i & 1
            if done[k] is not None:
                done[k].synchronize()  # the buffer's previous DMA has finished
            m = min(chunk, n - off)
            np.copyto(bufs[k].numpy()[:m], src[off:off + m], casting="unsafe")
            dst[off:off + m].copy_(bufs[k][:m], non_blocking=True)
            done[k] = torch.cuda.Event()
            done[k].record()
        for e in done:  # the buffers are free again when this returns
            if e is not None:
                e.synchronize()
    return out


def _d2h_staged(t) -> np.ndarray:
    """float32 NumPy copy of a (large) device tensor through pinned staging."""
    torch = _torch()
    src = t.detach().to(torch.float32).contiguous().view(-1)
    out = np.empty(tuple(t.shape), dtype=np.float32)
    dst = out.reshape(-1)
    n, chunk = src.numel(), _STAGE_BYTES // 4
    with _stage_lock:
        bufs = _staging()
        pending = [None, None]  # (event, offset, count) of the chunk each buffer holds

        def drain(k):
            e, o, m = pending[k]
            e.synchronize()
            dst[o:o + m] = bufs[k].numpy()[:m]
            pending[k] = None

        for i, off in enumerate(range(0, n, chunk)):
            k = i & 1
            if pending[k] is not None:
                drain(k)
            m = min(chunk, n - off)
            bufs[k][:m].copy_(src[off:off + m], non_blocking=True)
            e = torch.cuda.Event()
            e.record()
            pending[k] = (e, off, m)
        for k in sorted((k for k in (0, 1) if pending[k] is not None), key=lambda k: pending[k][1]):
            drain(k)
    return out


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
    if arr.size * 4 >= _STAGE_MIN_BYTES:
        return _h2d_staged(arr)
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
    if t.device.type == "cuda" and t.numel() * 4 >= _STAGE_MIN_BYTES:
        return _d2h_staged(t).astype(dt, copy=False)
    return t.cpu().numpy().astype(dt, copy=False)


def to_device_thw(a, name: str = "array"):
    """Frame-major (T,H,W) float32 device copy of an (H,W,T) array.

    NumPy arrays whose memory is already frame-major (an (H,W,T) view of
    (T,H,W) data) are uploaded as they lie; Fortran-ordered ones (as
    ``scipy.io.loadmat`` returns MATLAB arrays) are uploaded as (T,W,H) and
    each frame is transposed on the device.  Either way the host makes no
    full-size copy."""
    if not is_torch(a):
        arr = np.asarray(a)
        if arr.ndim == 3 and arr.dtype.kind in "fiub":
            thw = arr.transpose(2, 0, 1)
            if thw.flags.c_contiguous:
                return to_device(thw, name)
            twh = arr.transpose(2, 1, 0)
            if twh.flags.c_contiguous:
                return to_device(twh, name).transpose(1, 2).contiguous()
    return hwt_to_thw(to_device(a, name))


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
