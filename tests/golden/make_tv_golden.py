"""Write golden vectors for the Chambolle TV step and small GAP/ADMM-TV solves.

Run with the interpreter that has scikit-image 0.18.3 (the third-party
library the reference imports at PnP_SCI/python/pnp_sci_algo.py:5):

    /opt/conda/bin/python3.9 tests/golden/make_tv_golden.py

Nothing from the reference checkout is imported.  The TV results come from
``skimage.restoration.denoise_tv_chambolle`` itself; the solver loops around
it (GAP at pnp_sci_algo.py:636-650, ADMM at :805-836) are written out below
so that the fixtures record what that library returns inside them.  Outputs:

    tests/golden/tv_chambolle_golden.npz
    tests/golden/gaptv_small_golden.npz
    tests/golden/ssim_golden.npz
"""

import os
import sys

import numpy as np
import skimage
from skimage.restoration import denoise_tv_chambolle

HERE = os.path.dirname(os.path.abspath(__file__))


def smooth_frame(rng, h, w, scale):
    """Low-frequency frame: sum of a few random cosines (no filtering needed)."""
    r = np.arange(h)[:, None]
    c = np.arange(w)[None, :]
    img = np.zeros((h, w))
    for _ in range(3):
        fr, fc = rng.uniform(0, scale, 2)
        ph = rng.uniform(0, 2 * np.pi)
        img += np.cos(2 * np.pi * (fr * r / h + fc * c / w) + ph)
    return 0.5 + 0.1 * img


def tv_cases():
    rng = np.random.default_rng(1234)
    cases = []
    shapes = [(1, 1, 1), (2, 2, 1), (1, 7, 2), (6, 1, 2), (3, 5, 2), (5, 4, 3),
              (16, 16, 2)]
    for shape in shapes:
        for dt in (np.float32, np.float64):
            for w in (0.05, 0.3):
                for n in (1, 2, 3, 5):
                    img = rng.random(shape).astype(dt)
                    cases.append((img, w, n, 2e-4))
    # longer runs where the energy exit fires on smooth frames
    for dt in (np.float32, np.float64):
        img = np.stack([smooth_frame(rng, 48, 40, 1.5) for _ in range(3)], -1).astype(dt)
        img[..., 1] += 0.05 * rng.standard_normal((48, 40)).astype(dt)
        for n in (5, 30, 200):
            cases.append((img.copy(), 0.1, n, 2e-4))
        cases.append((img.copy(), 0.1, 30, 5e-2))
    # constant frame (E0 = 0, never exits) and a frame with one bright pixel
    for dt in (np.float32, np.float64):
        img = np.full((9, 11, 2), 0.25, dtype=dt)
        img[4, 5, 1] = 1.0
        cases.append((img, 0.2, 5, 2e-4))
    # ragged sizes of the kind the kernels tile (odd, prime, tall, wide)
    for shape in ((33, 17, 3), (40, 64, 2), (67, 131, 2), (131, 37, 2), (8, 200, 1)):
        img = rng.random(shape).astype(np.float32)
        cases.append((img, 0.1, 5, 2e-4))
    return cases


def write_tv():
    arrays = {}
    meta = []
    for k, (img, w, n, eps) in enumerate(tv_cases()):
        out = denoise_tv_chambolle(img, weight=w, eps=eps, n_iter_max=n,
                                   multichannel=True)
        assert out.dtype == img.dtype
        arrays['in_%d' % k] = img
        arrays['out_%d' % k] = out
        meta.append((w, n, eps))
    arrays['meta'] = np.array(meta, dtype=np.float64)
    arrays['skimage_version'] = np.array(skimage.__version__)
    np.savez_compressed(os.path.join(HERE, 'tv_chambolle_golden.npz'), **arrays)
    print('tv cases:', len(meta))


def gap_loop(y, mask, msum, x0, lam, accelerate, iters, w, n):
    """GAP outer loop (pnp_sci_algo.py:636-650) around scikit-image's TV."""
    x = x0
    y1 = np.zeros_like(y)
    for _ in range(iters):
        yb = np.sum(x * mask, axis=2)
        if accelerate:
            y1 = y1 + (y - yb)
            x = x + lam * (np.repeat(((y1 - yb) / msum)[:, :, None], mask.shape[2], 2) * mask)
        else:
            x = x + lam * (np.repeat(((y - yb) / msum)[:, :, None], mask.shape[2], 2) * mask)
        x = denoise_tv_chambolle(x, w, n_iter_max=n, multichannel=True)
    return x


def admm_loop(y, mask, msum, x0, lam, gamma, iters, w, n, clip):
    """ADMM outer loop (pnp_sci_algo.py:800-836; clip per joint_pnp_sci_algo.py:633)."""
    theta = x0
    b = np.zeros_like(x0)
    x = x0
    for _ in range(iters):
        yb = np.sum((theta + b) * mask, axis=2)
        x = (theta + b) + lam * (np.repeat(((y - yb) / (msum + gamma))[:, :, None],
                                           mask.shape[2], 2) * mask)
        theta = denoise_tv_chambolle(x - b, w, n_iter_max=n, multichannel=True)
        if clip:
            theta = np.clip(theta, 0, 1)
        b = b - (x - theta)
    return x


def write_gap():
    rng = np.random.default_rng(99)
    arrays = {}
    meta = []
    k = 0
    for (h, wd, t) in ((24, 20, 4), (33, 29, 6)):
        for dt in (np.float32, np.float64):
            truth = np.clip(np.stack([smooth_frame(rng, h, wd, 2.0) for _ in range(t)], -1), 0, 1)
            mask = (rng.random((h, wd, t)) < 0.5).astype(dt)
            truth = truth.astype(dt)
            y = np.sum(truth * mask, axis=2)
            msum = np.sum(mask, axis=2)
            msum[msum == 0] = 1
            x0 = np.repeat(y[:, :, None], t, 2) * mask
            runs = [('gap_acc', gap_loop(y, mask, msum, x0, 1.0, True, 12, 0.1, 5)),
                    ('gap_plain', gap_loop(y, mask, msum, x0, 1.0, False, 12, 0.1, 5)),
                    ('admm', admm_loop(y, mask, msum, x0, 1.0, 0.01, 12, 0.1, 5, False)),
                    ('admm_clip', admm_loop(y, mask, msum, x0, 1.0, 0.0, 12, 0.1, 5, True))]
            arrays['y_%d' % k] = y
            arrays['mask_%d' % k] = mask
            for name, res in runs:
                arrays['%s_%d' % (name, k)] = res
            meta.append((h, wd, t))
            k += 1
    arrays['meta'] = np.array(meta)
    np.savez_compressed(os.path.join(HERE, 'gaptv_small_golden.npz'), **arrays)
    print('gap cases:', len(meta))


def legacy_A(f, Phi):
    """Two-argument forward model the legacy solvers are called with."""
    return np.sum(f * Phi, axis=2)


def legacy_At(y, Phi):
    return np.multiply(np.repeat(y[:, :, np.newaxis], Phi.shape[2], axis=2), Phi)


def legacy_gap(y, Phi, Phi_sum, maxiter, step_size, weight, row, col):
    """GAP_TV_rec (pnp_sci_algo.py:866-882) around scikit-image's TV, progress
    printing left out: 30 TV iterations per call, y1 accumulated from zero."""
    y1 = np.zeros((row, col))
    f = legacy_At(y, Phi)
    for _ in range(maxiter):
        fb = legacy_A(f, Phi)
        y1 = y1 + (y - fb)
        f = f + np.multiply(step_size, legacy_At(np.divide(y1 - fb, Phi_sum), Phi))
        f = denoise_tv_chambolle(f, weight, n_iter_max=30, multichannel=True)
    return f


def legacy_admm(y, Phi, Phi_sum, maxiter, step_size, weight, row, col, ColT, eta):
    """ADMM_TV_rec (pnp_sci_algo.py:884-907) around scikit-image's TV: the TV
    weight and eta decay by 0.999 and 0.998 after every iteration, and the
    returned image is v (the last projection)."""
    theta = legacy_At(y, Phi)
    v = theta
    b = np.zeros((row, col, ColT))
    for _ in range(maxiter):
        yb = legacy_A(theta + b, Phi)
        v = (theta + b) + np.multiply(step_size, legacy_At(np.divide(y - yb, Phi_sum + eta), Phi))
        theta = denoise_tv_chambolle(v - b, weight, n_iter_max=30, multichannel=True)
        b = b - (v - theta)
        weight = 0.999 * weight
        eta = 0.998 * eta
    return v


def write_legacy():
    """Legacy two-argument solvers with scikit-image's 30-iteration TV inside.

    The TV weight is large enough that most frames run all 30 iterations and
    some stop early on the energy test, so the fixtures record both branches;
    ADMM runs long enough (25 iterations) for the decays to move the weight
    by 2.5% and eta by 5%."""
    rng = np.random.default_rng(2024)
    arrays = {}
    meta = []
    k = 0
    for (h, wd, t, iters, w, eta) in ((20, 24, 4, 6, 0.15, 0.01), (17, 31, 5, 25, 0.3, 0.05)):
        for dt in (np.float32, np.float64):
            truth = np.clip(np.stack([smooth_frame(rng, h, wd, 2.0) for _ in range(t)], -1), 0, 1)
            truth += 0.05 * rng.standard_normal(truth.shape)
            mask = (rng.random((h, wd, t)) < 0.5).astype(dt)
            truth = truth.astype(dt)
            y = np.sum(truth * mask, axis=2)
            msum = np.sum(mask, axis=2)
            msum[msum == 0] = 1
            arrays['y_%d' % k] = y
            arrays['mask_%d' % k] = mask
            arrays['gap_%d' % k] = legacy_gap(y, mask, msum, iters, 1.0, w, h, wd)
            arrays['admm_%d' % k] = legacy_admm(y, mask, msum, iters, 1.0, w, h, wd, t, eta)
            meta.append((h, wd, t, iters, w, eta))
            k += 1
    arrays['meta'] = np.array(meta, dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, 'legacy_golden.npz'), **arrays)
    print('legacy cases:', len(meta))


def write_psnr():
    """peak_signal_noise_ratio(ref, img, data_range=1.) as called at
    pnp_sci_algo.py:702, on float32 and float64 pairs (scikit-image casts
    both to float64 before the error)."""
    from skimage.metrics import peak_signal_noise_ratio
    rng = np.random.default_rng(11)
    arrays = {}
    vals = []
    cases = [((16, 16), 0.1, np.float64), ((16, 16), 0.1, np.float32),
             ((7, 9), 1e-3, np.float32), ((40, 33), 0.3, np.float64),
             ((5, 6), 1e-6, np.float64), ((64, 64), 0.02, np.float32)]
    for k, (shape, sigma, dt) in enumerate(cases):
        ref = rng.random(shape).astype(dt)
        img = (ref + sigma * rng.standard_normal(shape)).astype(dt)
        arrays['ref_%d' % k] = ref
        arrays['img_%d' % k] = img
        vals.append(peak_signal_noise_ratio(ref, img, data_range=1.))
    arrays['psnr'] = np.array(vals)
    np.savez_compressed(os.path.join(HERE, 'psnr_golden.npz'), **arrays)
    print('psnr cases:', len(vals))


def write_ssim():
    """structural_similarity(X, Y, data_range=1.) as called at pnp_sci_algo.py:703."""
    from skimage.metrics import structural_similarity
    rng = np.random.default_rng(7)
    arrays = {}
    vals = []
    for k, shape in enumerate(((7, 7), (9, 13), (32, 24), (64, 48))):
        X = rng.random(shape)
        Y = np.clip(X + 0.1 * rng.standard_normal(shape), 0, 1).astype(np.float32)
        arrays['x_%d' % k] = X
        arrays['y_%d' % k] = Y
        vals.append(structural_similarity(X, Y, data_range=1.))
    arrays['ssim'] = np.array(vals)
    np.savez_compressed(os.path.join(HERE, 'ssim_golden.npz'), **arrays)
    print('ssim cases:', len(vals))


if __name__ == '__main__':
    if skimage.__version__ != '0.18.3':
        sys.exit('expected scikit-image 0.18.3, found %s' % skimage.__version__)
    which = sys.argv[1:] or ['tv', 'gap', 'ssim', 'legacy', 'psnr']
    for name in which:
        {'tv': write_tv, 'gap': write_gap, 'ssim': write_ssim, 'legacy': write_legacy,
         'psnr': write_psnr}[name]()
