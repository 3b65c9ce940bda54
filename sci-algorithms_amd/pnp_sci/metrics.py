"""Per-frame image-quality metrics on the device.

PSNR (``compare_psnr``) and SSIM (``compare_ssim``) as the reference reports
them per frame at the end of a solve (PnP_SCI/python/pnp_sci_algo.py:699-705,
857-863), both from scikit-image 0.18.3 with ``data_range=1``.
"""

from __future__ import annotations

import math

from . import _lib
from .device import sq_error

__all__ = ["psnr_frames", "ssim_frames"]


def psnr_frames(truth, x, data_range: float = 1.0):
    """10*log10(data_range^2/mse) per frame of two (T,H,W) device tensors."""
    T, H, W = x.shape
    out = []
    for s in sq_error(truth, x, per_frame=True).cpu().tolist():
        mse = s / (H * W)
        out.append(math.inf if mse == 0 else 10 * math.log10(data_range ** 2 / mse))
    return out


def ssim_frames(truth, x, data_range: float = 1.0):
    """Mean SSIM per frame of two (T,H,W) device tensors (scigap_ssim).
    Frames smaller than the 7x7 window raise scikit-image's ValueError."""
    import torch
    T, H, W = x.shape
    if min(H, W) < 7:
        raise ValueError("win_size exceeds image extent.  If the input is a multichannel (color) image, set multichannel=True.")
    out = torch.empty(T, dtype=torch.float64, device=x.device)
    L = _lib.lib()
    _lib.check(L.scigap_ssim(truth.data_ptr(), x.data_ptr(), out.data_ptr(), H, W, T,
                             float(data_range), _lib.stream_ptr()))
    return out.cpu().tolist()
