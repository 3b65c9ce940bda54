"""GAP-TV / ADMM-TV reconstruction for video SCI on MI355X.

Drop-in mirror of the TV path of ``PnP_SCI/python/pnp_sci_algo.py``:
``gap_denoise`` (:536-706), ``admm_denoise`` (:708-864),
``gap_denoise_bayer`` (:20-265), ``admmdenoise_cacti`` (:479-534) and the legacy ``GAP_TV_rec`` /
``ADMM_TV_rec`` (:866-907), plus the clipped ADMM of
``joint_pnp_sci_algo.py:502-663`` as ``admm_denoise(..., clip=True)``.
Same argument names, defaults, return tuples and error behaviour; every outer
iteration runs as fused HIP kernels (scigap_solve) in fp32.

Differences a caller can observe:
  * Arithmetic is fp32 whatever the input dtype; results are returned in the
    input's dtype and container (NumPy or torch).
  * ``denoiser`` must be 'tv'.  The deep and wavelet denoisers of the
    reference are not part of this path and raise NotImplementedError.
  * ``multichannel=False`` (3-D TV over (H,W,T)) is not implemented.
  * The operators ``A``/``At`` must be the SCI mask operators (any callables
    equal to ``A_(x, Phi)`` / ``At_(y, Phi)``): the mask is recovered once as
    ``At(ones)`` and checked against ``A`` on a random probe.
"""

from __future__ import annotations

import math
import time

import numpy as np

from . import _lib
from .device import hwt_to_thw, is_torch, thw_to_hwt, to_device, to_device_thw, to_host_like
from .metrics import psnr_frames, ssim_frames

__all__ = ["gap_denoise", "admm_denoise", "admmdenoise_cacti", "GAP_TV_rec", "ADMM_TV_rec",
           "gap_denoise_bayer", "recover_mask", "solve_device"]

_OTHER_DENOISERS = ("wavelet", "ffdnet", "fastdvdnet", "vnlnet")


def _torch():
    import torch
    return torch


# ---------------------------------------------------------------------------
# Operator recovery
# ---------------------------------------------------------------------------

def recover_mask(A, At, y, check: bool = True):
    """Device (T,H,W) mask behind the opaque operators A, At.

    ``At_(y, Phi)`` is ``y*Phi`` per frame, so ``At(ones) = Phi`` exactly
    (utils.py:17-26).  ``A`` is then checked on a random probe against
    ``sum_t x*Phi``; a mismatch means the callables are not SCI operators.
    """
    phi = getattr(At, "scigap_phi", None)
    if phi is not None:
        return phi
    torch = _torch()
    ones = torch.ones_like(y) if is_torch(y) else np.ones(np.shape(y), dtype=np.asarray(y).dtype)
    mask_hwt = At(ones)
    if len(np.shape(mask_hwt)) != 3 or tuple(np.shape(mask_hwt)[:2]) != tuple(np.shape(y)):
        raise ValueError("At(ones) must return an (H,W,T) mask, got shape %s"
                         % (tuple(np.shape(mask_hwt)),))
    phi = to_device_thw(mask_hwt, "At(ones)")
    if check:
        H, W, T = np.shape(mask_hwt)
        rng = np.random.default_rng(12345)
        probe = rng.random((H, W, T), dtype=np.float32)
        if is_torch(y):
            probe_in = torch.from_numpy(probe).to(device=y.device, dtype=y.dtype)
        else:
            probe_in = probe.astype(np.asarray(y).dtype, copy=False)
        got = to_device(A(probe_in), "A(probe)")
        pd = to_device_thw(probe, "probe")
        want = torch.empty((H, W), dtype=torch.float32, device=pd.device)
        L = _lib.lib()
        _lib.check(L.scigap_forward(pd.data_ptr(), phi.data_ptr(), want.data_ptr(), H, W, T,
                                    _lib.stream_ptr()))
        scale = float(want.abs().max().item()) + 1e-30
        if got.shape != want.shape or float((got - want).abs().max().item()) > 1e-4 * scale:
            raise ValueError("A and At are not the SCI forward model and its transpose "
                             "(A(x) != sum_t x*At(1)); only A_(x,Phi)/At_(y,Phi) are supported")
    return phi


# ---------------------------------------------------------------------------
# Device solve
# ---------------------------------------------------------------------------

def solve_device(y, phi, phi_sum, *, method="gap", iters=50, _lambda=1.0, accelerate=True,
                 gamma=0.01, clip=False, tv_weight=0.1, tv_iter_max=5, tv_eps=2e-4,
                 x0=None, truth=None, tv_weight_decay=1.0, gamma_decay=1.0, stream=None):
    """One GAP/ADMM-TV solve on device tensors already in library layout.

    y, phi_sum: (H,W) float32; phi, x0, truth: (T,H,W) float32, all on the
    device.  Returns (x (T,H,W), sq_err (iters,) fp64 or None,
    early_exits (iters,) int32).
    """
    torch = _torch()
    T, H, W = phi.shape
    p = _lib.Params()
    L = _lib.lib()
    L.scigap_default_params(p, H, W, T)
    p.method = _lib.METHOD_ADMM if method == "admm" else _lib.METHOD_GAP
    p.accelerate = 1 if accelerate else 0
    p.clip = 1 if clip else 0
    p.iters = int(iters)
    p.lambda_ = float(_lambda)
    p.gamma = float(gamma)
    p.tv_weight = float(tv_weight)
    p.tv_eps = float(tv_eps)
    p.tv_iter_max = int(tv_iter_max)
    p.tv_weight_decay = float(tv_weight_decay)
    p.gamma_decay = float(gamma_decay)
    nbytes = L.scigap_solve_workspace_bytes(p)
    if nbytes == 0:
        raise _lib.ScigapError(-1, L.scigap_last_error().decode(errors="replace"))
    dev = phi.device
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    x = torch.empty((T, H, W), dtype=torch.float32, device=dev)
    sq = torch.empty(max(iters, 1), dtype=torch.float64, device=dev) if truth is not None else None
    stats = torch.empty(max(iters, 1), dtype=torch.int32, device=dev)
    ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
    _lib.check(L.scigap_solve(p, ptr(y), ptr(phi), ptr(phi_sum), ptr(x0), ptr(x), ptr(truth),
                              ptr(sq), ptr(stats), ptr(ws), nbytes, _lib.stream_ptr(stream)))
    return x, (sq[:iters] if sq is not None else None), stats[:iters]


# ---------------------------------------------------------------------------
# Reference-shaped entry points
# ---------------------------------------------------------------------------

def _check_denoiser(denoiser, multichannel):
    d = denoiser.lower()
    if d in _OTHER_DENOISERS:
        raise NotImplementedError("denoiser %r is outside the GAP/ADMM-TV path" % denoiser)
    if d != "tv":
        raise ValueError('Unsupported denoiser {}!'.format(denoiser))
    if not multichannel:
        raise NotImplementedError("multichannel=False (3-D TV) is not implemented")


def _stages(sigma, iter_max):
    """Noise-level stages (pnp_sci_algo.py:628-631); for TV they only add up."""
    if not isinstance(sigma, list):
        sigma = [sigma]
    if not isinstance(iter_max, list):
        iter_max = [iter_max] * len(sigma)
    return sigma, iter_max


def _psnr_from_sq(sq, n):
    out = []
    for s in sq:
        mse = s / n
        out.append(100 if mse == 0 else 20 * math.log10(1.0 / math.sqrt(mse)))
    return out


def _print_progress(tag, psnr_all, sigma, iter_max):
    """Progress lines of pnp_sci_algo.py:683-696 (every fifth iteration)."""
    k = 0
    for idx, nsig in enumerate(sigma):
        for _ in range(iter_max[idx]):
            if (k + 1) % 5 == 0 and k < len(psnr_all):
                if nsig is not None:
                    if nsig < 1:
                        print('  {0}-TV iteration {1: 3d}, sigma {2: 3g}/255, '
                              'PSNR {3:2.2f} dB.'.format(tag, k + 1, nsig * 255, psnr_all[k]))
                    else:
                        print('  {0}-TV iteration {1: 3d}, sigma {2: 3g}, '
                              'PSNR {3:2.2f} dB.'.format(tag, k + 1, nsig, psnr_all[k]))
                else:
                    print('  {0}-TV iteration {1: 3d}, PSNR {2:2.2f} dB.'.format(
                        tag, k + 1, psnr_all[k]))
            k += 1


def _run(method, y, Phi_sum, A, At, *, _lambda, accelerate, gamma, clip, denoiser, iter_max,
         sigma, tv_weight, tv_iter_max, multichannel, x0, X_orig, show_iqa, tvm="tv_chambolle",
         tv_weight_decay=1.0, gamma_decay=1.0, tv_eps=2e-4):
    _check_denoiser(denoiser, multichannel)
    if tvm in ("ITV3D_FGP", "ITV2D_cham") and method == "gap":
        # pnp_sci_algo.py:651-654 call functions that are never defined.
        raise NameError("name of the TV denoiser for tvm=%r is not defined" % tvm)
    if tvm != "tv_chambolle" and method == "gap":
        # pnp_sci_algo.py:655-658: TypeError raised and printed, TV skipped.
        print("Exception: ", repr(TypeError("no such tv denoiser")))
        tv_iter_max = 1  # identity denoiser
    sigma, iter_max = _stages(sigma, iter_max)
    iters = int(sum(iter_max))
    phi = recover_mask(A, At, y)
    T, H, W = phi.shape
    yd = to_device(y, "y")
    psd = to_device(Phi_sum, "Phi_sum")
    if tuple(yd.shape) != (H, W) or tuple(psd.shape) != (H, W):
        raise ValueError("y and Phi_sum must be (H,W) = (%d,%d)" % (H, W))
    x0d = to_device_thw(x0, "x0") if x0 is not None else None
    want_iqa = show_iqa and X_orig is not None
    truth = to_device_thw(X_orig, "X_orig") if X_orig is not None else None
    x, sq, _ = solve_device(yd, phi, psd, method=method, iters=iters, _lambda=_lambda,
                            accelerate=accelerate, gamma=gamma, clip=clip, tv_weight=tv_weight,
                            tv_iter_max=tv_iter_max, tv_eps=tv_eps, x0=x0d,
                            truth=truth if want_iqa else None,
                            tv_weight_decay=tv_weight_decay, gamma_decay=gamma_decay)
    psnr_all = []
    if want_iqa:
        psnr_all = _psnr_from_sq(sq.cpu().tolist(), T * H * W)
        _print_progress(method.upper(), psnr_all, sigma, iter_max)
    psnr_, ssim_ = [], []
    if X_orig is not None:
        psnr_ = psnr_frames(truth, x)
        ssim_ = ssim_frames(truth, x)  # raises for frames under 7x7, as the reference does
    return to_host_like(thw_to_hwt(x), y), psnr_, ssim_, psnr_all


def gap_denoise(y, Phi_sum, A, At, _lambda=1, accelerate=True,
                denoiser='tv', iter_max=50, noise_estimate=False, sigma=None,
                tv_weight=0.1, tv_iter_max=5, multichannel=True, x0=None,
                X_orig=None, model=None, show_iqa=True, tvm='tv_chambolle'):
    """GAP-TV (pnp_sci_algo.py:536-706).  Returns (x, psnr_, ssim_, psnr_all)."""
    return _run("gap", y, Phi_sum, A, At, _lambda=_lambda, accelerate=accelerate, gamma=0.0,
                clip=False, denoiser=denoiser, iter_max=iter_max, sigma=sigma,
                tv_weight=tv_weight, tv_iter_max=tv_iter_max, multichannel=multichannel,
                x0=x0, X_orig=X_orig, show_iqa=show_iqa, tvm=tvm)


def admm_denoise(y, Phi_sum, A, At, _lambda=1, gamma=0.01,
                 denoiser='tv', iter_max=50, noise_estimate=False, sigma=None,
                 tv_weight=0.1, tv_iter_max=5, multichannel=True, x0=None, model=None,
                 X_orig=None, show_iqa=True, clip=False):
    """ADMM-TV (pnp_sci_algo.py:708-864).  ``clip=True`` is the variant of
    joint_pnp_sci_algo.py:502-663 (theta clipped to [0,1] after TV; that
    function defaults gamma to 0).  Returns (x, psnr_, ssim_, psnr_all)."""
    return _run("admm", y, Phi_sum, A, At, _lambda=_lambda, accelerate=False, gamma=gamma,
                clip=clip, denoiser=denoiser, iter_max=iter_max, sigma=sigma,
                tv_weight=tv_weight, tv_iter_max=tv_iter_max, multichannel=multichannel,
                x0=x0, X_orig=X_orig, show_iqa=show_iqa)


_BAYER = [[0, 0], [0, 1], [1, 0], [1, 1]]  # 2x2 sub-grids (pnp_sci_algo.py:100-102)


def gap_denoise_bayer(y_bayer, Phi_bayer, _lambda=1, accelerate=True, denoiser='tv',
                      iter_max=50, noise_estimate=True, sigma=None, tv_weight=0.1,
                      tv_iter_max=5, multichannel=True, x0_bayer=None, X_orig=None,
                      model=None, show_iqa=True):
    """GAP-TV on a Bayer-mosaic measurement (pnp_sci_algo.py:20-265).

    The reference splits the mosaic into its four 2x2 sub-grids, applies the
    GAP update to each and one multichannel TV to all nmask*4 sub-grid frames
    per outer iteration.  The TV is per frame, so the four sub-grids are
    independent GAP-TV solves: each runs as one fused solve here, and the
    results are written back into the mosaic.  y_bayer (H,W), Phi_bayer
    (H,W,nmask) with H and W even.  Returns (x_bayer, psnr_, ssim_, psnr_all),
    psnr_all on the whole mosaic per iteration.
    """
    torch = _torch()
    _check_denoiser(denoiser, multichannel)
    sigma, iter_max = _stages(sigma, iter_max)
    iters = int(sum(iter_max))
    yd = to_device(y_bayer, "y_bayer")
    phid = to_device(Phi_bayer, "Phi_bayer")
    nrow, ncol, nmask = phid.shape
    if nrow % 2 or ncol % 2 or tuple(yd.shape) != (nrow, ncol):
        raise ValueError("y_bayer (H,W) and Phi_bayer (H,W,nmask) need equal, even H and W")
    want_iqa = show_iqa and X_orig is not None
    truth = to_device(X_orig, "X_orig") if X_orig is not None else None
    x0d = to_device(x0_bayer, "x0_bayer") if x0_bayer is not None else None
    L = _lib.lib()
    x_bayer = torch.empty((nrow, ncol, nmask), dtype=torch.float32, device=phid.device)
    sq_total = None
    for b0, b1 in _BAYER:
        y_c = yd[b0::2, b1::2].contiguous()
        phi_c = hwt_to_thw(phid[b0::2, b1::2].contiguous())
        T, H, W = phi_c.shape
        ps_c = torch.empty((H, W), dtype=torch.float32, device=phid.device)
        _lib.check(L.scigap_phi_sum(phi_c.data_ptr(), ps_c.data_ptr(), H, W, T,
                                    _lib.stream_ptr()))
        x0_c = hwt_to_thw(x0d[b0::2, b1::2].contiguous()) if x0d is not None else None
        tr_c = hwt_to_thw(truth[b0::2, b1::2].contiguous()) if want_iqa else None
        x_c, sq, _ = solve_device(y_c, phi_c, ps_c, method="gap", iters=iters, _lambda=_lambda,
                                  accelerate=accelerate, tv_weight=tv_weight,
                                  tv_iter_max=tv_iter_max, x0=x0_c, truth=tr_c)
        x_bayer[b0::2, b1::2] = thw_to_hwt(x_c)
        if sq is not None:
            sq_total = sq if sq_total is None else sq_total + sq
    psnr_all = []
    if want_iqa:
        psnr_all = _psnr_from_sq(sq_total.cpu().tolist(), nrow * ncol * nmask)
        stages = [None] * len(sigma) if noise_estimate else sigma
        _print_progress("GAP", psnr_all, stages, iter_max)
    psnr_, ssim_ = [], []
    if X_orig is not None:
        xt = hwt_to_thw(x_bayer)
        tt = hwt_to_thw(truth)
        psnr_ = psnr_frames(tt, xt)
        ssim_ = ssim_frames(tt, xt)
    return to_host_like(x_bayer, Phi_bayer), psnr_, ssim_, psnr_all


def admmdenoise_cacti(meas, mask, A, At, projmeth='admm', v0=None, orig=None,
                      iframe=0, nframe=1, MAXB=1., maskdirection='plain', **args):
    """Per-coded-frame driver (pnp_sci_algo.py:479-534).

    Divides measurement and truth by MAXB, computes Phi_sum once with the
    zero fix (:491-492, on the device), flips warm starts and results for
    'updown'/'downup' masks, and returns (x_, t_, psnr_, ssim_, psnrall_)
    with x_ float32 (H,W,nmask*nframe) and t_ the wall time in seconds.
    """
    torch = _torch()
    nmask = mask.shape[-1]
    md = to_device(mask, "mask")
    phi = hwt_to_thw(md)
    T, H, W = phi.shape
    msum = torch.empty((H, W), dtype=torch.float32, device=phi.device)
    L = _lib.lib()
    _lib.check(L.scigap_phi_sum(phi.data_ptr(), msum.data_ptr(), H, W, T, _lib.stream_ptr()))
    if getattr(At, "scigap_phi", None) is None:
        # Probe the caller's operators once, then reuse the recovered mask.
        phi = recover_mask(A, At, meas[..., iframe] if not is_torch(meas) else meas[..., iframe])

    class _Bound:
        scigap_phi = phi

        def __call__(self, v):  # pragma: no cover - never called by the solvers
            raise RuntimeError("bound operator is not callable")

    A_b, At_b = _Bound(), _Bound()
    x_ = np.zeros((*mask.shape[:-1], nmask * nframe), dtype=np.float32)
    psnr_, ssim_, psnrall_ = [], [], []
    begin = time.time()
    t_ = 0.0
    msum_host = msum.cpu().numpy()
    for kf in range(nframe):
        print('%s-%s Reconstruction coded frame block %2d of %2d ...'
              % (projmeth.upper(), args['denoiser'].upper(), kf + 1, nframe))
        orig_k = None
        if orig is not None:
            orig_k = np.asarray(orig[..., (kf + iframe) * nmask:(kf + iframe + 1) * nmask]) / MAXB
        meas_k = np.asarray(meas[..., kf + iframe]) / MAXB
        flip = ((maskdirection.lower() == 'updown' and (kf + iframe) % 2 == 1) or
                (maskdirection.lower() == 'downup' and (kf + iframe) % 2 == 0))
        v0_k = None
        if v0 is not None:
            v0_k = np.asarray(v0[:, :, kf * nmask:(kf + 1) * nmask])
            if flip:
                v0_k = v0_k[..., ::-1]
        if projmeth.lower() == 'admm':
            res = admm_denoise(meas_k, msum_host, A_b, At_b, x0=v0_k, X_orig=orig_k, **args)
        elif projmeth.lower() == 'gap':
            res = gap_denoise(meas_k, msum_host, A_b, At_b, x0=v0_k, X_orig=orig_k, **args)
        else:
            # The reference prints this and then reads the unbound result of
            # the solve (pnp_sci_algo.py:518-529): UnboundLocalError.
            print('Unsupported projection method %s' % projmeth.upper())
            raise UnboundLocalError("local variable 'x_k' referenced before assignment")
        x_k, psnr_k, ssim_k, psnrall_k = res
        if flip:
            x_k = x_k[..., ::-1]
            psnr_k = psnr_k[::-1]
            ssim_k = ssim_k[::-1]
            psnrall_k = psnrall_k[::-1]
        t_ = time.time() - begin
        x_[..., kf * nmask:(kf + 1) * nmask] = x_k
        psnr_.extend(psnr_k)
        ssim_.extend(ssim_k)
        psnrall_.append(psnrall_k)
    return x_, t_, psnr_, ssim_, psnrall_


def GAP_TV_rec(y, Phi, A, At, Phi_sum, maxiter, step_size, weight, row, col, ColT, X_ori):
    """Legacy accelerated GAP-TV with 30 TV iterations per call
    (pnp_sci_algo.py:866-882); ``A(f, Phi)``/``At(y, Phi)`` are not called,
    the mask ``Phi`` is used directly.  Progress printing is omitted."""
    x, _, _, _ = _run("gap", y, Phi_sum, _MaskOnly(Phi), _MaskOnly(Phi), _lambda=step_size,
                      accelerate=True, gamma=0.0, clip=False, denoiser='tv', iter_max=maxiter,
                      sigma=None, tv_weight=weight, tv_iter_max=30, multichannel=True, x0=None,
                      X_orig=None, show_iqa=False)
    return x


def ADMM_TV_rec(y, Phi, A, At, Phi_sum, maxiter, step_size, weight, row, col, ColT, eta, X_ori):
    """Legacy ADMM-TV (pnp_sci_algo.py:884-907): 30 TV iterations per call,
    TV weight x0.999 and eta x0.998 after every iteration."""
    x, _, _, _ = _run("admm", y, Phi_sum, _MaskOnly(Phi), _MaskOnly(Phi), _lambda=step_size,
                      accelerate=False, gamma=eta, clip=False, denoiser='tv', iter_max=maxiter,
                      sigma=None, tv_weight=weight, tv_iter_max=30, multichannel=True, x0=None,
                      X_orig=None, show_iqa=False, tv_weight_decay=0.999, gamma_decay=0.998)
    return x


class _MaskOnly:
    """Operator stand-in carrying an explicit (H,W,T) mask."""

    def __init__(self, Phi):
        self.scigap_phi = to_device_thw(Phi, "Phi")

    def __call__(self, v):  # pragma: no cover
        raise RuntimeError("mask-only operator is not callable")
