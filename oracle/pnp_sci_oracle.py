"""CPU oracle for the PnP-SCI GAP-TV / ADMM-TV path -- TEST INFRASTRUCTURE ONLY.

This module restates, in NumPy, the arithmetic of the reference path so that
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
can check and time the HIP implementation against it.  Nothing in the product
(``sci-algorithms_amd/``) imports it; the product fails loudly when its HIP
library is missing instead of falling back here.

Every function works in the dtype it is given (float32 or float64), as the
reference does: the reference drivers call the solvers with float32 arrays
(``PnP_SCI/python/pnp_sci_demo_kobe.py:55-57``) and also with float64 ones
(``PnP_SCI/python/pnp_sci_test_orig.py:115``).  Running it on float64 inputs is
the FP64 oracle.

Sources restated (paths relative to the reference checkout):

* ``PnP_SCI/python/utils.py:10-36``         -- ``A_``, ``At_``, ``psnr``
* ``PnP_SCI/python/pnp_sci_algo.py:20-265``  -- ``gap_denoise_bayer`` (TV)
* ``PnP_SCI/python/pnp_sci_algo.py:479-534`` -- ``admmdenoise_cacti``
* ``PnP_SCI/python/pnp_sci_algo.py:536-706`` -- ``gap_denoise``
* ``PnP_SCI/python/pnp_sci_algo.py:708-864`` -- ``admm_denoise``
* ``PnP_SCI/python/pnp_sci_algo.py:866-907`` -- ``GAP_TV_rec``, ``ADMM_TV_rec``
* ``PnP_SCI/python/joint_pnp_sci_algo.py:502-663`` -- ADMM with ``clip(theta,0,1)``
* scikit-image 0.18.3 (third-party, not in the reference; imported at
  ``pnp_sci_algo.py:5``) ``skimage/restoration/_denoise.py:315-393``
  (``_denoise_tv_chambolle_nd``) and ``:396-479`` (``denoise_tv_chambolle``),
  and ``skimage/metrics/simple_metrics.py`` ``peak_signal_noise_ratio``.

How each function is pinned (the tests under ``tests/`` do the checking;
fixtures under ``tests/golden/`` are written by scikit-image 0.18.3 itself,
``make_tv_golden.py``):

* ``A_``/``At_``/``phi_sum``: closed forms; adjointness <A x, y> = <x, At y>,
  At(1) = Phi, A(At(y)) = y * sum_t Phi_t^2, a zero-mask pixel gets sum 1.
* ``tv_chambolle_nd``/``denoise_tv_chambolle``: bit-exact float32 and float64
  golden vectors (127 cases), plus the invariants n_iter_max=1 -> identity,
  constant frame unchanged, frame mean preserved, the w -> inf one-update
  limit img + Laplacian/4 (scipy.ndimage.laplace), and the early-exit branch.
* ``gap_denoise``/``admm_denoise``/``admm_denoise_clip``: bit-exact golden
  solves (the reference's loops around scikit-image's TV), plus: plain GAP
  with an identity denoiser is an exact projection (A x = y where
  Phi_sum > 0); the first accelerated step equals the plain step; ADMM with
  gamma=0 and an identity denoiser reproduces plain GAP.
* ``GAP_TV_rec``/``ADMM_TV_rec``: bit-exact golden solves with scikit-image's
  30-iteration TV and the 0.999/0.998 decays (``legacy_golden.npz``); the
  fixtures are checked to contain early TV exits and to reject a missing or
  misplaced decay; GAP_TV_rec also meets ``gap_denoise`` on the float64 ones.
* ``admmdenoise_cacti``: every block equals the direct solve, with the warm
  start and the result (and its PSNR/SSIM lists) reversed on the odd ('updown')
  or even ('downup') coded frames counted from ``iframe``, and MAXB scaling.
* ``gap_denoise_bayer``: equals four independent ``gap_denoise`` runs on the
  2x2 sub-grids reassembled into the mosaic (the TV is per frame, so the
  joint multichannel TV decouples); mosaic PSNRs equal the closed form.
* ``psnr``: scikit-image ``peak_signal_noise_ratio`` values with data_range 1
  (``psnr_golden.npz``; equal in float64), closed forms (20 dB at mse 0.01,
  40 dB at 1e-4), 100 for a zero error (``utils.py:28-36``).
* ``compare_psnr``: the same scikit-image values to 1e-12 in float32 and
  float64, inf for a zero error, and the data_range scaling.
* ``compare_ssim``: golden values written by scikit-image 0.18.3.
"""

from __future__ import annotations

import math

import numpy as np

__all__ = [
    "A_", "At_", "psnr", "compare_psnr", "phi_sum", "tv_chambolle_nd",
    "denoise_tv_chambolle", "gap_denoise", "admm_denoise", "admm_denoise_clip",
    "admmdenoise_cacti", "GAP_TV_rec", "ADMM_TV_rec", "compare_ssim", "gap_denoise_bayer",
]


# ---------------------------------------------------------------------------
# Forward model (PnP_SCI/python/utils.py:10-26)
# ---------------------------------------------------------------------------

def A_(x, Phi):
    """Snapshot forward model y[h,w] = sum_t x[h,w,t]*Phi[h,w,t] (utils.py:10-15)."""
    return (x * Phi).sum(axis=2)


def At_(y, Phi):
    """Adjoint x[h,w,t] = y[h,w]*Phi[h,w,t] (utils.py:17-26).

    The reference materialises ``np.repeat(y[:,:,None], T, axis=2)`` before the
    multiply; the values are identical to a broadcast, and the repeat is kept
    here so that the timed CPU baseline follows the reference's memory traffic.
    """
    return np.repeat(y[:, :, np.newaxis], Phi.shape[2], axis=2) * Phi


def phi_sum(Phi):
    """Mask sum over frames with zero pixels set to 1 (pnp_sci_algo.py:491-492)."""
    s = Phi.sum(axis=tuple(range(2, Phi.ndim)))
    s[s == 0] = 1
    return s


def psnr(ref, img):
    """PSNR with peak 1 and 100 dB for an exact match (utils.py:28-36)."""
    mse = np.mean((ref - img) ** 2)
    if mse == 0:
        return 100
    return 20 * math.log10(1.0 / math.sqrt(mse))


def compare_psnr(ref, img, data_range=1.0):
    """scikit-image ``peak_signal_noise_ratio`` as called at pnp_sci_algo.py:702.

    10*log10(data_range^2 / mse).  As scikit-image 0.18.3 does
    (``metrics/simple_metrics.py`` ``_as_floats`` and ``mean_squared_error``),
    both images are promoted to ``result_type(ref, img, float32)``, the
    difference is squared in that type and the mean is accumulated in
    float64; inf for a zero error.
    """
    ref, img = np.asarray(ref), np.asarray(img)
    ft = np.result_type(ref.dtype, img.dtype, np.float32)
    d = np.asarray(ref, ft) - np.asarray(img, ft)
    err = np.mean(d ** 2, dtype=np.float64)
    if err == 0:
        return math.inf
    return 10 * math.log10(data_range ** 2 / err)


def compare_ssim(X, Y, data_range=1.0):
    """scikit-image 0.18.3 ``structural_similarity`` for 2-D images with its
    defaults, as called at pnp_sci_algo.py:703 (metrics/_structural_similarity.py):
    7x7 uniform window (scipy.ndimage.uniform_filter), sample covariance
    (49/48), K1=0.01, K2=0.03, float64, mean over the image cropped by 3."""
    from scipy.ndimage import uniform_filter
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    if min(X.shape) < 7:
        raise ValueError("win_size exceeds image extent.  If the input is a multichannel (color) image, set multichannel=True.")
    win, npix = 7, 49.0
    cov = npix / (npix - 1.0)
    ux, uy = uniform_filter(X, win), uniform_filter(Y, win)
    uxx, uyy, uxy = uniform_filter(X * X, win), uniform_filter(Y * Y, win), uniform_filter(X * Y, win)
    vx, vy, vxy = cov * (uxx - ux * ux), cov * (uyy - uy * uy), cov * (uxy - ux * uy)
    c1, c2 = (0.01 * data_range) ** 2, (0.03 * data_range) ** 2
    S = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux ** 2 + uy ** 2 + c1) * (vx + vy + c2))
    return float(S[3:-3, 3:-3].mean())


# ---------------------------------------------------------------------------
# Chambolle TV (scikit-image 0.18.3 restoration/_denoise.py:315-479)
# ---------------------------------------------------------------------------

def tv_chambolle_nd(image, weight=0.1, eps=2.e-4, n_iter_max=200, _stats=None):
    """Chambolle's dual projection for ROF denoising of one n-D image.

    Follows ``_denoise_tv_chambolle_nd`` (``_denoise.py:315-393``):

    * the dual field p (one component per axis) starts at zero on every call;
    * iteration i>0 forms d = -div-like term: d = -sum_ax p_ax, then adds
      p_ax shifted by +1 along ax (so the first index along ax gets no
      contribution), and ``out = image + d``; iteration 0 uses ``out = image``
      and d = 0;
    * g_ax is the forward difference of ``out`` along ax, zero in the last
      slice; |g| is the Euclidean norm over ax;
    * the energy E = (sum d^2 + weight * sum |g|) / image.size; in float32 the
      two sums are float32 and the rest is float64 (NumPy 1.x promotes
      ``weight * np.float32``), which only matters for the exit test;
    * p <- (p - tau*g) / (1 + (tau/weight)*|g|) with tau = 1/(2*ndim);
    * exit when i>0 and |E_prev - E| < eps*E_0, returning the ``out`` formed at
      the start of that iteration.  Without an exit the result is the ``out``
      of iteration n_iter_max-1, i.e. after n_iter_max-1 dual updates.

    ``_stats`` (a dict, test hook) receives ``returned``: the iteration whose
    ``out`` is returned, and ``margin``: the smallest distance of an exit test
    from its threshold, min |(|E_prev - E| / (eps*E_0)) - 1| over the tests
    taken (inf when none was), which tells a decision that rounding could flip.
    """
    nd = image.ndim
    dt = image.dtype
    # The reference passes a Python float; a NumPy float64 scalar would make
    # ``mag *= tau/weight`` compute in float64 under NumPy 2 promotion rules.
    weight = float(weight)
    tau = 1.0 / (2.0 * nd)
    p = np.zeros((nd,) + image.shape, dtype=dt)
    grad = np.zeros_like(p)
    d = np.zeros_like(image)
    out = image
    e_first = e_prev = 0.0
    margin = float("inf")
    it = 0
    while it < n_iter_max:
        if it > 0:
            d = -p.sum(0)
            for ax in range(nd):
                dst = [slice(None)] * nd
                src = [slice(None)] * nd
                dst[ax] = slice(1, None)
                src[ax] = slice(0, -1)
                d[tuple(dst)] += p[ax][tuple(src)]
            out = image + d
        energy = float((d ** 2).sum())
        for ax in range(nd):
            lead = [slice(None)] * nd
            lead[ax] = slice(0, -1)
            grad[ax][tuple(lead)] = np.diff(out, axis=ax)
        mag = np.sqrt((grad ** 2).sum(axis=0))
        energy += weight * float(mag.sum())
        mag *= tau / weight
        mag += 1.0
        p -= tau * grad
        p /= mag[np.newaxis]
        energy /= float(image.size)
        if it == 0:
            e_first = e_prev = energy
        else:
            if _stats is not None and eps * e_first > 0:
                margin = min(margin, abs(abs(e_prev - energy) / (eps * e_first) - 1.0))
            if abs(e_prev - energy) < eps * e_first:
                break
            e_prev = energy
        it += 1
    if _stats is not None:
        _stats["returned"] = min(it, n_iter_max - 1)
        _stats["margin"] = margin
    return out


def denoise_tv_chambolle(image, weight=0.1, eps=2.e-4, n_iter_max=200,
                         multichannel=False):
    """Wrapper of ``_denoise.py:396-479``: per-channel (last axis) when
    ``multichannel``, else one n-D problem.  Non-float input is converted to
    float64 in [0,1] the way ``img_as_float`` does for unsigned integers."""
    image = np.asarray(image)
    if image.dtype.kind != 'f':
        if image.dtype.kind == 'u':
            image = image.astype(np.float64) / np.iinfo(image.dtype).max
        else:
            raise TypeError("only float or unsigned images are restated")
    if multichannel:
        out = np.zeros_like(image)
        for c in range(image.shape[-1]):
            out[..., c] = tv_chambolle_nd(image[..., c], weight, eps, n_iter_max)
        return out
    return tv_chambolle_nd(image, weight, eps, n_iter_max)


# ---------------------------------------------------------------------------
# PnP solvers (PnP_SCI/python/pnp_sci_algo.py)
# ---------------------------------------------------------------------------

def _tv_dispatch(x, tv_weight, tv_iter_max, multichannel, tvm, eps=2.e-4):
    """TV branch of pnp_sci_algo.py:647-658.

    'tv_chambolle' runs scikit-image's Chambolle TV.  'ITV3D_FGP' and
    'ITV2D_cham' name functions the reference never defines, so they raise
    NameError there (:651-654).  Any other name raises TypeError inside a
    try block whose handler prints it, leaving x untouched (:655-658).
    """
    if tvm == 'tv_chambolle':
        return denoise_tv_chambolle(x, tv_weight, eps=eps, n_iter_max=tv_iter_max,
                                    multichannel=multichannel)
    if tvm in ('ITV3D_FGP', 'ITV2D_cham'):
        raise NameError("name of the TV denoiser for tvm=%r is not defined" % tvm)
    print("Exception: ", repr(TypeError("no such tv denoiser")))
    return x


def _check_denoiser(denoiser):
    if denoiser.lower() != 'tv':
        raise ValueError('Unsupported denoiser {}!'.format(denoiser))


def _stages(sigma, iter_max):
    """Noise-level stages of pnp_sci_algo.py:628-631 (lists run back to back)."""
    if not isinstance(sigma, list):
        sigma = [sigma]
    if not isinstance(iter_max, list):
        iter_max = [iter_max] * len(sigma)
    return sigma, iter_max


def _final_metrics(x, X_orig):
    """Per-frame PSNR and SSIM of pnp_sci_algo.py:699-705."""
    if X_orig is None:
        return [], []
    psnr_ = [compare_psnr(X_orig[..., m], x[..., m]) for m in range(x.shape[-1])]
    # the reference calls compare_ssim for every frame (so frames smaller
    # than the 7x7 window raise scikit-image's ValueError there)
    ssim_ = [compare_ssim(X_orig[..., m], x[..., m]) for m in range(x.shape[-1])]
    return psnr_, ssim_


def gap_denoise(y, Phi_sum, A, At, _lambda=1, accelerate=True,
                denoiser='tv', iter_max=50, noise_estimate=False, sigma=None,
                tv_weight=0.1, tv_iter_max=5, multichannel=True, x0=None,
                X_orig=None, model=None, show_iqa=True, tvm='tv_chambolle', _tv_eps=2.e-4):
    """GAP-TV, pnp_sci_algo.py:536-706 with denoiser='tv'.

    ``_tv_eps`` (test hook, not in the reference) sets the TV exit tolerance.

    Returns ``(x, psnr_, ssim_, psnr_all)``.
    """
    _check_denoiser(denoiser)
    if x0 is None:
        x0 = At(y)
    sigma, iter_max = _stages(sigma, iter_max)
    y1 = np.zeros_like(y)
    x = x0
    psnr_all = []
    for idx, _nsig in enumerate(sigma):
        for _ in range(iter_max[idx]):
            yb = A(x)
            if accelerate:
                y1 = y1 + (y - yb)
                x = x + _lambda * (At((y1 - yb) / Phi_sum))
            else:
                x = x + _lambda * (At((y - yb) / Phi_sum))
            x = _tv_dispatch(x, tv_weight, tv_iter_max, multichannel, tvm, _tv_eps)
            if show_iqa and X_orig is not None:
                psnr_all.append(psnr(X_orig, x))
    return (x, *_final_metrics(x, X_orig), psnr_all)


def admm_denoise(y, Phi_sum, A, At, _lambda=1, gamma=0.01,
                 denoiser='tv', iter_max=50, noise_estimate=False, sigma=None,
                 tv_weight=0.1, tv_iter_max=5, multichannel=True, x0=None,
                 model=None, X_orig=None, show_iqa=True, clip=False, _tv_eps=2.e-4):
    """ADMM-TV, pnp_sci_algo.py:708-864 with denoiser='tv'.

    ``clip=True`` gives the variant of joint_pnp_sci_algo.py:502-663, which
    clips theta to [0,1] after the denoiser (:633) and defaults gamma to 0.
    The returned image is x (the projection output), as in the reference.
    """
    _check_denoiser(denoiser)
    if x0 is None:
        x0 = At(y)
    sigma, iter_max = _stages(sigma, iter_max)
    x = x0
    theta = x0
    b = np.zeros_like(x0)
    psnr_all = []
    for idx, _nsig in enumerate(sigma):
        for _ in range(iter_max[idx]):
            yb = A(theta + b)
            x = (theta + b) + _lambda * (At((y - yb) / (Phi_sum + gamma)))
            theta = denoise_tv_chambolle(x - b, tv_weight, eps=_tv_eps, n_iter_max=tv_iter_max,
                                         multichannel=multichannel)
            if clip:
                theta = np.clip(theta, 0, 1)
            b = b - (x - theta)
            if show_iqa and X_orig is not None:
                psnr_all.append(psnr(X_orig, x))
    return (x, *_final_metrics(x, X_orig), psnr_all)


def admm_denoise_clip(y, Phi_sum, A, At, _lambda=1, gamma=0.0, **kw):
    """joint_pnp_sci_algo.py:502-663: ADMM-TV with theta clipped to [0,1]."""
    return admm_denoise(y, Phi_sum, A, At, _lambda=_lambda, gamma=gamma,
                        clip=True, **kw)


_BAYER = [[0, 0], [0, 1], [1, 0], [1, 1]]  # 2x2 sub-grids, pnp_sci_algo.py:100-102


def gap_denoise_bayer(y_bayer, Phi_bayer, _lambda=1, accelerate=True, denoiser='tv',
                      iter_max=50, noise_estimate=True, sigma=None, tv_weight=0.1,
                      tv_iter_max=5, multichannel=True, x0_bayer=None, X_orig=None,
                      model=None, show_iqa=True, _tv_eps=2.e-4):
    """GAP-TV on a Bayer mosaic, pnp_sci_algo.py:20-265 with denoiser='tv'.

    The mosaic is split into its four 2x2 sub-grids (:104-136); every outer
    iteration applies the GAP update to each sub-grid (:150-157) and then one
    multichannel TV over the nmask*4 sub-grid frames (:161-165); the result is
    reassembled into the mosaic (:248-250).  Per-iteration PSNR is taken on
    the mosaic (:227-231).  The reference allocates the sub-grid arrays in
    float32 (:108-111); here they follow the input, so float64 inputs give
    the FP64 oracle.  Returns ``(x_bayer, psnr_, ssim_, psnr_all)``.
    """
    _check_denoiser(denoiser)
    sigma, iter_max = _stages(sigma, iter_max)
    nrow, ncol, nmask = Phi_bayer.shape
    dt = np.float32
    yall = np.zeros([nrow // 2, ncol // 2, 4], dtype=dt)
    Phiall = np.zeros([nrow // 2, ncol // 2, nmask, 4], dtype=dt)
    Phi_sumall = np.zeros([nrow // 2, ncol // 2, 4], dtype=dt)
    x0all = np.zeros([nrow // 2, ncol // 2, nmask, 4], dtype=dt)
    if np.asarray(y_bayer).dtype == np.float64:
        dt = np.float64
        yall, Phiall, Phi_sumall, x0all = (a.astype(dt) for a in (yall, Phiall, Phi_sumall, x0all))
    for ib, b in enumerate(_BAYER):
        yall[..., ib] = y_bayer[b[0]::2, b[1]::2]
        Phiall[..., ib] = Phi_bayer[b[0]::2, b[1]::2]
        ps = np.sum(Phiall[..., ib], axis=2)
        ps[ps == 0] = 1
        Phi_sumall[..., ib] = ps
        if x0_bayer is None:
            x0all[..., ib] = At_(yall[..., ib], Phiall[..., ib])
        else:
            x0all[..., ib] = x0_bayer[b[0]::2, b[1]::2]
    y1all = np.zeros_like(yall)
    xall = x0all
    x_bayer = np.zeros_like(Phi_bayer)
    psnr_all = []
    for idx, _nsig in enumerate(sigma):
        for _ in range(iter_max[idx]):
            for ib in range(4):
                yb = A_(xall[..., ib], Phiall[..., ib])
                if accelerate:
                    y1all[..., ib] += (yall[..., ib] - yb)
                    xall[..., ib] += _lambda * At_((y1all[..., ib] - yb) / Phi_sumall[..., ib],
                                                   Phiall[..., ib])
                else:
                    xall[..., ib] += _lambda * At_((yall[..., ib] - yb) / Phi_sumall[..., ib],
                                                   Phiall[..., ib])
            xv = xall.reshape([nrow // 2, ncol // 2, nmask * 4])
            xv = denoise_tv_chambolle(xv, tv_weight, eps=_tv_eps, n_iter_max=tv_iter_max,
                                      multichannel=multichannel)
            xall = xv.reshape([nrow // 2, ncol // 2, nmask, 4])
            if show_iqa and X_orig is not None:
                for ib, b in enumerate(_BAYER):
                    x_bayer[b[0]::2, b[1]::2] = xall[..., ib]
                psnr_all.append(compare_psnr(X_orig, x_bayer, data_range=1.))
    for ib, b in enumerate(_BAYER):
        x_bayer[b[0]::2, b[1]::2] = xall[..., ib]
    psnr_, ssim_ = [], []
    if X_orig is not None:
        for m in range(nmask):
            psnr_.append(compare_psnr(X_orig[:, :, m], x_bayer[:, :, m], data_range=1.))
            ssim_.append(compare_ssim(X_orig[:, :, m], x_bayer[:, :, m], data_range=1.))
    return x_bayer, psnr_, ssim_, psnr_all


def admmdenoise_cacti(meas, mask, A, At, projmeth='admm', v0=None, orig=None,
                      iframe=0, nframe=1, MAXB=1., maskdirection='plain',
                      **args):
    """Per-coded-frame driver, pnp_sci_algo.py:479-534 (timing omitted: t_=0).

    Divides each measurement (and the truth) by MAXB, computes Phi_sum once,
    flips the warm start and the result along frames for 'updown'/'downup'
    masks on alternate coded frames, and stacks the results as float32.
    """
    nmask = mask.shape[-1]
    msum = phi_sum(mask)
    x_ = np.zeros((*mask.shape[:-1], nmask * nframe), dtype=np.float32)
    psnr_, ssim_, psnrall_ = [], [], []
    for kf in range(nframe):
        orig_k = None
        if orig is not None:
            orig_k = orig[..., (kf + iframe) * nmask:(kf + iframe + 1) * nmask] / MAXB
        meas_k = meas[..., kf + iframe] / MAXB
        flip = ((maskdirection.lower() == 'updown' and (kf + iframe) % 2 == 1) or
                (maskdirection.lower() == 'downup' and (kf + iframe) % 2 == 0))
        v0_k = None
        if v0 is not None:
            v0_k = v0[:, :, kf * nmask:(kf + 1) * nmask]
            if flip:
                v0_k = v0_k[..., ::-1]
        if projmeth.lower() == 'admm':
            res = admm_denoise(meas_k, msum, A, At, x0=v0_k, X_orig=orig_k, **args)
        elif projmeth.lower() == 'gap':
            res = gap_denoise(meas_k, msum, A, At, x0=v0_k, X_orig=orig_k, **args)
        else:
            # pnp_sci_algo.py:518-529: printed, then the unbound x_k is read
            print('Unsupported projection method %s' % projmeth.upper())
            raise UnboundLocalError("local variable 'x_k' referenced before assignment")
        x_k, psnr_k, ssim_k, psnrall_k = res
        if flip:
            x_k = x_k[..., ::-1]
            psnr_k = psnr_k[::-1]
            ssim_k = ssim_k[::-1]
            psnrall_k = psnrall_k[::-1]
        x_[..., kf * nmask:(kf + 1) * nmask] = x_k
        psnr_.extend(psnr_k)
        ssim_.extend(ssim_k)
        psnrall_.append(psnrall_k)
    return x_, 0.0, psnr_, ssim_, psnrall_


# ---------------------------------------------------------------------------
# Legacy two-argument solvers (pnp_sci_algo.py:866-907)
# ---------------------------------------------------------------------------

def GAP_TV_rec(y, Phi, A, At, Phi_sum, maxiter, step_size, weight, row, col,
               ColT, X_ori):
    """Accelerated GAP-TV with 30 TV iterations per call (pnp_sci_algo.py:866-882).

    ``A(f, Phi)``/``At(y, Phi)`` take the mask explicitly; progress printing
    is omitted."""
    y1 = np.zeros((row, col))
    f = At(y, Phi)
    for _ in range(maxiter):
        fb = A(f, Phi)
        y1 = y1 + (y - fb)
        f = f + np.multiply(step_size, At(np.divide(y1 - fb, Phi_sum), Phi))
        f = denoise_tv_chambolle(f, weight, n_iter_max=30, multichannel=True)
    return f


def ADMM_TV_rec(y, Phi, A, At, Phi_sum, maxiter, step_size, weight, row, col,
                ColT, eta, X_ori):
    """ADMM-TV with per-iteration decay of the TV weight (x0.999) and of eta
    (x0.998), 30 TV iterations per call (pnp_sci_algo.py:884-907)."""
    theta = At(y, Phi)
    v = theta
    b = np.zeros((row, col, ColT))
    for _ in range(maxiter):
        yb = A(theta + b, Phi)
        v = (theta + b) + np.multiply(step_size, At(np.divide(y - yb, Phi_sum + eta), Phi))
        theta = denoise_tv_chambolle(v - b, weight, n_iter_max=30, multichannel=True)
        b = b - (v - theta)
        weight = 0.999 * weight
        eta = 0.998 * eta
    return v
