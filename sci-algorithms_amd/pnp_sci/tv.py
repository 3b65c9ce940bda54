"""Chambolle TV denoising on the device with scikit-image's signature.

``denoise_tv_chambolle(image, weight=0.1, eps=2e-4, n_iter_max=200,
multichannel=False)`` as in scikit-image 0.18.3
(skimage/restoration/_denoise.py:396-479), which the reference calls at
PnP_SCI/python/pnp_sci_algo.py:650 and :812 with ``multichannel=True``.
Here only the per-channel 2-D form is built: ``multichannel=True`` on an
(H,W,C) array, or a single (H,W) image.
"""

from __future__ import annotations

from . import _lib
from .device import hwt_to_thw, thw_to_hwt, to_device, to_host_like

__all__ = ["denoise_tv_chambolle", "tv_device"]


def tv_device(image_thw, weight=0.1, eps=2e-4, n_iter_max=200, return_stop=False, depth=0):
    """TV of a (T,H,W) float32 device tensor; returns a new tensor (and the
    per-frame iteration whose image was returned, if asked).  ``depth`` is
    the number of dual updates fused per launch (1..6; 0 = automatic), a
    tuning knob that does not change the result."""
    import torch
    T, H, W = image_thw.shape
    L = _lib.lib()
    nbytes = L.scigap_tv_workspace_bytes_depth(H, W, T, int(n_iter_max), int(depth))
    if nbytes == 0:
        raise ValueError("invalid TV arguments (n_iter_max must be >= 1, depth in [0, 6])")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=image_thw.device)
    out = torch.empty_like(image_thw)
    stop = torch.empty(T, dtype=torch.int32, device=image_thw.device) if return_stop else None
    _lib.check(L.scigap_tv_chambolle_depth(image_thw.data_ptr(), out.data_ptr(), H, W, T,
                                           float(weight), float(eps), int(n_iter_max),
                                           int(depth),
                                           stop.data_ptr() if stop is not None else None,
                                           ws.data_ptr(), nbytes, _lib.stream_ptr()))
    return (out, stop) if return_stop else out


def denoise_tv_chambolle(image, weight=0.1, eps=2.e-4, n_iter_max=200, multichannel=False):
    """scikit-image's Chambolle TV for 2-D frames on the HIP device."""
    img = to_device(image, "image")
    if img.dim() == 2:
        out = tv_device(img[None], weight, eps, n_iter_max)[0]
        return to_host_like(out, image)
    if img.dim() == 3 and multichannel:
        out = tv_device(hwt_to_thw(img), weight, eps, n_iter_max)
        return to_host_like(thw_to_hwt(out), image)
    raise NotImplementedError("only 2-D images or multichannel=True (H,W,C) stacks are supported")
