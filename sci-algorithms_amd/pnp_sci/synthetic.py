"""Seeded synthetic CACTI data for the benchmark and parity workloads.

The reference's datasets (kobe32 and the bm_256_10f videos) are Git-LFS
pointers in the checkout, so every workload here is generated from a seed:

* truth video (H,W,T) in [0,1]: a smooth background -- uniform noise
  low-passed with a Gaussian of sigma = min(H,W)/32 (applied in the Fourier
  domain with periodic boundaries) and rescaled to [0, 0.5] -- plus 5
  Gaussian blobs (sigma ~ U(8,20) px, amplitude ~ U(0.3,0.7), start ~ uniform
  over the frame, heading ~ U(0, 2pi), speed ~ U(1,3) px/frame) moving
  linearly; clipped to [0,1];
* masks: i.i.d. Bernoulli(p) per pixel and frame, or one Bernoulli(0.5) base
  mask shifted down by t rows for frame t (the shifting-mask CACTI pattern);
* measurement y = sum_t Phi_t * x_t (utils.py:10-15), optionally plus
  Gaussian noise of standard deviation noise_rel * max(y).

Draw order from ``numpy.random.default_rng(seed)``: background noise, the 5
blobs (sigma, amplitude, row, column, heading, speed each), masks, noise.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Problem", "truth_video", "make_problem", "problem_256x256x8", "problem_2048x2048x20",
           "mixed_queue_sizes", "mixed_queue_problem", "unrolled_batch4_256x256x8",
           "unrolled_batch2_1024x1024x20"]


@dataclass
class Problem:
    truth_thw: np.ndarray  # (T,H,W) float32, contiguous
    mask_thw: np.ndarray   # (T,H,W) float32, contiguous
    y: np.ndarray          # (H,W) float32
    phi_sum: np.ndarray    # (H,W) float32, zeros replaced by 1

    @property
    def truth(self) -> np.ndarray:
        """(H,W,T) view in the reference's layout."""
        return self.truth_thw.transpose(1, 2, 0)

    @property
    def mask(self) -> np.ndarray:
        """(H,W,T) view in the reference's layout."""
        return self.mask_thw.transpose(1, 2, 0)


def _lowpass(noise: np.ndarray, sigma: float) -> np.ndarray:
    H, W = noise.shape
    fy = np.fft.fftfreq(H)[:, None]
    fx = np.fft.rfftfreq(W)[None, :]
    g = np.exp(-2.0 * (np.pi ** 2) * (sigma ** 2) * (fy ** 2 + fx ** 2))
    return np.fft.irfft2(np.fft.rfft2(noise) * g, s=(H, W))


def truth_video(rng: np.random.Generator, H: int, W: int, T: int, nblobs: int = 5,
                frame_major: bool = False) -> np.ndarray:
    """(H,W,T) video (or (T,H,W) with frame_major) in [0,1], float32."""
    bg = _lowpass(rng.random((H, W)), min(H, W) / 32.0)
    lo, hi = bg.min(), bg.max()
    bg = 0.5 * (bg - lo) / (hi - lo if hi > lo else 1.0)
    params = []
    for _ in range(nblobs):
        sig = rng.uniform(8, 20)
        amp = rng.uniform(0.3, 0.7)
        r0 = rng.uniform(0, H)
        c0 = rng.uniform(0, W)
        head = rng.uniform(0, 2 * np.pi)
        speed = rng.uniform(1, 3)
        params.append((sig, amp, r0, c0, head, speed))
    rows = np.arange(H, dtype=np.float64)
    cols = np.arange(W, dtype=np.float64)
    video = np.empty((T, H, W), dtype=np.float32)
    for t in range(T):
        frame = bg.copy()
        for sig, amp, r0, c0, head, speed in params:
            rc = r0 + t * speed * np.sin(head)
            cc = c0 + t * speed * np.cos(head)
            gr = np.exp(-((rows - rc) ** 2) / (2 * sig * sig))
            gc = np.exp(-((cols - cc) ** 2) / (2 * sig * sig))
            frame += np.outer(amp * gr, gc)
        np.clip(frame, 0.0, 1.0, out=frame)
        video[t] = frame
    return video if frame_major else video.transpose(1, 2, 0)


def make_problem(H: int, W: int, T: int, *, seed: int = 0, mask: str = "bernoulli",
                 p: float = 0.5, noise_rel: float = 0.0, rng=None) -> Problem:
    rng = rng if rng is not None else np.random.default_rng(seed)
    truth = truth_video(rng, H, W, T, frame_major=True)
    if mask == "bernoulli":
        # drawn in (H,W,T) order so the mask of a pixel's frames is contiguous in the stream
        phi = np.ascontiguousarray((rng.random((H, W, T)) < p).astype(np.float32).transpose(2, 0, 1))
    elif mask == "shift":
        base = (rng.random((H, W)) < p).astype(np.float32)
        phi = np.stack([np.roll(base, t, axis=0) for t in range(T)], axis=0)
    else:
        raise ValueError("unknown mask kind %r" % mask)
    y = np.einsum("thw,thw->hw", truth, phi, dtype=np.float64)
    if noise_rel > 0:
        y = y + rng.normal(0.0, noise_rel * y.max(), size=y.shape)
    y = y.astype(np.float32)
    ps = phi.sum(axis=0)
    ps[ps == 0] = 1
    return Problem(truth_thw=truth, mask_thw=phi, y=y, phi_sum=ps.astype(np.float32))


def problem_256x256x8(seed: int = 0) -> Problem:
    """256x256x8, i.i.d. Bernoulli(0.5) masks, noise-free."""
    return make_problem(256, 256, 8, seed=seed)


def problem_2048x2048x20(seed: int = 0, H: int = 2048, W: int = 2048, T: int = 20) -> Problem:
    """2048x2048x20, shifting Bernoulli(0.5) mask, noise sigma = 0.01*max(y)."""
    return make_problem(H, W, T, seed=seed, mask="shift", noise_rel=0.01)


CONFIG3_SIZES = ((256, 256, 8), (512, 512, 10), (1024, 1024, 20), (512, 768, 16))


def mixed_queue_sizes(seed: int = 0, count: int = 64):
    """Sizes and mask densities of the mixed-size queue (uniform draws)."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(CONFIG3_SIZES), size=count)
    ps = rng.uniform(0.3, 0.7, size=count)
    return [(CONFIG3_SIZES[i], float(p)) for i, p in zip(idx, ps)]


def mixed_queue_problem(k: int, size, p: float, seed: int = 0) -> Problem:
    H, W, T = size
    return make_problem(H, W, T, seed=seed * 1000 + k + 1, p=p)


def unrolled_batch(batch: int, H: int, W: int, T: int, seed: int = 0, mask: str = "logits"):
    """Training batch for the unrolled GAP-TV: truth videos (batch,T,H,W) and
    either mask logits ~ N(0,1) (T,H,W) (``mask="logits"``, phi = sigmoid) or
    a shifting binary Bernoulli(0.5) mask (``mask="shift"``), float32."""
    rng = np.random.default_rng(seed)
    truth = np.stack([truth_video(rng, H, W, T, frame_major=True) for _ in range(batch)])
    if mask == "logits":
        m = rng.standard_normal((T, H, W)).astype(np.float32)
    elif mask == "shift":
        base = (rng.random((H, W)) < 0.5).astype(np.float32)
        m = np.stack([np.roll(base, t, axis=0) for t in range(T)], axis=0)
    else:
        raise ValueError("unknown mask kind %r" % mask)
    return truth, m


def unrolled_batch4_256x256x8(seed: int = 0):
    """Unrolled GAP-TV training: batch 4 at 256x256x8, mask logits ~ N(0,1)
    (K=10 outer x 5 TV iterations)."""
    return unrolled_batch(4, 256, 256, 8, seed=seed, mask="logits")


def unrolled_batch2_1024x1024x20(seed: int = 0):
    """Large unrolled GAP-TV: batch 2 at 1024x1024x20 with a shifting binary
    mask (K=20 outer x 5 TV iterations, learnable lambda_k / tv_weight_k)."""
    return unrolled_batch(2, 1024, 1024, 20, seed=seed, mask="shift")
