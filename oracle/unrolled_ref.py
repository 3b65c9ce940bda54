"""Plain-PyTorch reference of the unrolled GAP-TV -- TEST INFRASTRUCTURE ONLY.

The differentiable model the HIP library implements with hand-written
backward kernels (``scigap_unrolled_forward``/``_backward``), written with
ordinary torch operations so that autograd supplies reference gradients.  Run
it in float64 (on the CPU, or on a GPU device as a reference computation).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` use it, as the
checker; the product never imports it.

Model, per sample b (shapes in the library's (T,H,W) layout):

    phi   = sigmoid(logits)  or the mask as given
    y_b   = sum_t phi_t * truth_{b,t}   or the measurement as given
    S     = sum_t phi_t     (exact zeros replaced by 1, pnp_sci_algo.py:491-492)
    x_0   = phi * y_b       (At(y), pnp_sci_algo.py:625-627),  y1 = 0
    for k < K:              (accelerated GAP, pnp_sci_algo.py:638-650)
        yb  = sum_t phi_t x_t
        y1  = y1 + (y - yb)
        x   = x + lambda_k * phi * (y1 - yb) / S
        x   = TV_{w_k}(x) with tv_iter_max - 1 Chambolle dual updates per
              frame (scikit-image 0.18.3 _denoise.py:341-393, no energy exit)

How it is pinned (tests/test_unrolled_ref.py):

* forward with constant lambda and weight equals the NumPy oracle's
  ``gap_denoise(..., _tv_eps=0)``, itself pinned to scikit-image goldens;
  per-iteration schedules equal the oracle's A_/At_/tv_chambolle_nd composed
  in the reference's loop order;
* autograd gradients equal FP64 central differences of the NumPy oracle's
  forward on small problems (mask logits, lambda_k, weight_k, y).
"""

from __future__ import annotations

import torch

TAU = 0.25  # 1/(2*ndim) for 2-D frames, _denoise.py:375


def tv_fixed(v: torch.Tensor, weight, n_iter_max: int) -> torch.Tensor:
    """Chambolle TV of every (H,W) frame of v (any leading dims) with exactly
    n_iter_max iterations of _denoise.py:341-393 (n_iter_max-1 dual updates;
    the last iteration only forms its output)."""
    p0 = torch.zeros_like(v)
    p1 = torch.zeros_like(v)
    out = v
    for i in range(n_iter_max):
        if i > 0:
            # d = -p.sum(0) + p shifted by one along each axis (_denoise.py:352-361)
            d = -(p0 + p1)
            d = d + torch.nn.functional.pad(p0[..., :-1, :], (0, 0, 1, 0))
            d = d + torch.nn.functional.pad(p1[..., :, :-1], (1, 0, 0, 0))
            out = v + d
        else:
            out = v
        if i == n_iter_max - 1:
            break
        g0 = torch.nn.functional.pad(out[..., 1:, :] - out[..., :-1, :], (0, 0, 0, 1))
        g1 = torch.nn.functional.pad(out[..., :, 1:] - out[..., :, :-1], (0, 1, 0, 0))
        # |g|, with a zero derivative where g = 0 (it is 0 on the last row and
        # column by construction)
        norm = torch.sqrt((g0 * g0 + g1 * g1).clamp_min(1e-300))
        den = 1.0 + norm * (TAU / weight)
        p0 = (p0 - TAU * g0) / den
        p1 = (p1 - TAU * g1) / den
    return out


def phi_sum(phi: torch.Tensor) -> torch.Tensor:
    s = phi.sum(0)
    return torch.where(s == 0, torch.ones_like(s), s)


def unrolled_gap_tv(phi: torch.Tensor, y: torch.Tensor, lambdas, weights,
                    tv_iter_max: int = 5, checkpoint: bool = False) -> torch.Tensor:
    """x_K for every sample.  phi (T,H,W), y (B,H,W), lambdas/weights (K,)
    tensors or sequences.  With checkpoint=True every outer iteration is
    wrapped in torch.utils.checkpoint (for large problems)."""
    K = len(lambdas)
    S = phi_sum(phi)
    x = phi.unsqueeze(0) * y.unsqueeze(1)
    y1 = torch.zeros_like(y)

    def step(x, y1, lam, w):
        yb = (phi.unsqueeze(0) * x).sum(1)
        y1 = y1 + (y - yb)
        x = x + lam * phi.unsqueeze(0) * ((y1 - yb) / S).unsqueeze(1)
        return tv_fixed(x, w, tv_iter_max), y1

    for k in range(K):
        if checkpoint:
            x, y1 = torch.utils.checkpoint.checkpoint(step, x, y1, lambdas[k], weights[k],
                                                      use_reentrant=False)
        else:
            x, y1 = step(x, y1, lambdas[k], weights[k])
    return x


def measure(phi: torch.Tensor, truth: torch.Tensor) -> torch.Tensor:
    """y_b = sum_t phi_t * truth_{b,t} (utils.py:10-15) for truth (B,T,H,W)."""
    return (phi.unsqueeze(0) * truth).sum(1)


def mse_loss(phi, truth, lambdas, weights, tv_iter_max=5, checkpoint=False):
    """Mean squared error of the reconstruction from y = A(truth)."""
    x = unrolled_gap_tv(phi, measure(phi, truth), lambdas, weights, tv_iter_max, checkpoint)
    return ((x - truth) ** 2).mean()
