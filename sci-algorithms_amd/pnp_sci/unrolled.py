"""Differentiable unrolled GAP-TV on the HIP library.

``unrolled_gap_tv`` is a ``torch.autograd.Function`` over
``scigap_unrolled_forward``/``scigap_unrolled_backward``: K outer iterations of
accelerated GAP-TV (PnP_SCI/python/pnp_sci_algo.py:625-650 with
denoiser='tv') with per-iteration lambda_k and tv_weight_k, each TV call
running exactly tv_iter_max-1 Chambolle dual updates (scikit-image 0.18.3
_denoise.py:341-393 without the energy early exit, which would make the map
discontinuous).  Gradients flow to the mask (or its logits, phi =
sigmoid(logits)), to lambda_k and tv_weight_k, and to the measurement y.

The forward pass keeps x_k and the residual r_k of every outer iteration (one
checkpoint per outer iteration); the backward pass reverses one outer
iteration at a time with the hand-written kernels, recomputing the inner TV
states in registers.

Tensors use the library's (T,H,W) frame-major layout: mask (T,H,W), truth
(B,T,H,W), y (B,H,W), output (B,T,H,W).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib

__all__ = ["unrolled_gap_tv", "UnrolledGAPTV", "unrolled_sizes"]


def _desc(H, W, T, batch, iters, tv_iter_max, logits, measure):
    d = _lib.Unrolled()
    d.H, d.W, d.T = H, W, T
    d.batch, d.iters, d.tv_iter_max = batch, iters, tv_iter_max
    d.logits, d.measure = int(bool(logits)), int(bool(measure))
    return d


def unrolled_sizes(H, W, T, batch, iters, tv_iter_max=5, logits=True, measure=True):
    """(checkpoint_bytes, workspace_bytes) of one forward/backward pair."""
    L = _lib.lib()
    d = _desc(H, W, T, batch, iters, tv_iter_max, logits, measure)
    ck = L.scigap_unrolled_checkpoint_bytes(ctypes.byref(d))
    ws = L.scigap_unrolled_workspace_bytes(ctypes.byref(d))
    if ck == 0 or ws == 0:
        raise _lib.ScigapError(-1, "unsupported unrolled configuration %r"
                               % ((H, W, T, batch, iters, tv_iter_max),))
    return ck, ws


def _host_floats(t: torch.Tensor):
    v = t.detach().to("cpu", torch.float32).contiguous()
    return v, ctypes.c_void_p(v.data_ptr())


class _UnrolledFn(torch.autograd.Function):

    @staticmethod
    def forward(ctx, mask, meas, lambdas, weights, tv_iter_max, logits, measure):
        if mask.dtype != torch.float32 or meas.dtype != torch.float32:
            raise TypeError("unrolled GAP-TV computes in float32")
        if not (mask.is_cuda and meas.is_cuda):
            raise ValueError("mask and measurements must be device tensors")
        mask = mask.contiguous()
        meas = meas.contiguous()
        T, H, W = mask.shape
        B = meas.shape[0]
        expect = (B, T, H, W) if measure else (B, H, W)
        if tuple(meas.shape) != expect:
            raise ValueError("measurement shape %s, expected %s" % (tuple(meas.shape), expect))
        K = lambdas.numel()
        if weights.numel() != K:
            raise ValueError("lambdas and tv weights need one entry per outer iteration")
        L = _lib.lib()
        d = _desc(H, W, T, B, K, tv_iter_max, logits, measure)
        ck_bytes = L.scigap_unrolled_checkpoint_bytes(ctypes.byref(d))
        ws_bytes = L.scigap_unrolled_workspace_bytes(ctypes.byref(d))
        if ck_bytes == 0 or ws_bytes == 0:
            _lib.check(-1 if ck_bytes == 0 else -4)
        lam_h, lam_p = _host_floats(lambdas)
        w_h, w_p = _host_floats(weights)
        x = torch.empty((B, T, H, W), dtype=torch.float32, device=mask.device)
        ckpt = torch.empty(ck_bytes, dtype=torch.uint8, device=mask.device)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=mask.device)
        _lib.check(L.scigap_unrolled_forward(ctypes.byref(d), lam_p, w_p, mask.data_ptr(),
                                             meas.data_ptr(), x.data_ptr(), ckpt.data_ptr(),
                                             ck_bytes, ws.data_ptr(), ws_bytes,
                                             _lib.stream_ptr()))
        del ws
        ctx.save_for_backward(mask, meas, lambdas, weights)
        ctx.ckpt = ckpt
        ctx.host = (lam_h, w_h)
        ctx.desc = (H, W, T, B, K, tv_iter_max, logits, measure)
        return x

    @staticmethod
    def backward(ctx, gx):
        if ctx.ckpt is None:
            raise RuntimeError("unrolled_gap_tv: the per-iteration checkpoints were released by "
                               "the first backward pass; call backward once per forward (a "
                               "second pass through this graph, e.g. after retain_graph=True, "
                               "needs a new forward)")
        mask, meas, lambdas, weights = ctx.saved_tensors
        H, W, T, B, K, n, logits, measure = ctx.desc
        L = _lib.lib()
        d = _desc(H, W, T, B, K, n, logits, measure)
        lam_h, w_h = ctx.host
        gx = gx.to(torch.float32).contiguous()
        need_mask, need_meas, need_lam, need_w = ctx.needs_input_grad[:4]
        g_mask = torch.empty_like(mask) if need_mask else None
        g_meas = torch.empty_like(meas) if (need_meas and not measure) else None
        g_lam = torch.empty(K, dtype=torch.float64, device=mask.device) if need_lam else None
        g_w = torch.empty(K, dtype=torch.float64, device=mask.device) if need_w else None
        ws_bytes = L.scigap_unrolled_workspace_bytes(ctypes.byref(d))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=mask.device)

        def ptr(t):
            return None if t is None else t.data_ptr()

        _lib.check(L.scigap_unrolled_backward(
            ctypes.byref(d), ctypes.c_void_p(lam_h.data_ptr()), ctypes.c_void_p(w_h.data_ptr()),
            mask.data_ptr(), meas.data_ptr(), ctx.ckpt.data_ptr(), ctx.ckpt.numel(),
            gx.data_ptr(), ptr(g_mask), ptr(g_meas), ptr(g_lam), ptr(g_w), ws.data_ptr(),
            ws_bytes, _lib.stream_ptr()))
        ctx.ckpt = None
        if g_lam is not None:
            g_lam = g_lam.to(device=lambdas.device, dtype=lambdas.dtype)
        if g_w is not None:
            g_w = g_w.to(device=weights.device, dtype=weights.dtype)
        return g_mask, g_meas, g_lam, g_w, None, None, None


def unrolled_gap_tv(mask, meas, lambdas, tv_weights, tv_iter_max=5, logits=True, measure=True):
    """x_K (B,T,H,W) of the unrolled GAP-TV.

    mask        (T,H,W) logits (``logits=True``, phi = sigmoid(mask)) or phi
    meas        truth videos (B,T,H,W) with ``measure=True`` (y = A(truth) is
                formed on the device), else measurements y (B,H,W)
    lambdas     (K,) step sizes lambda_k
    tv_weights  (K,) TV weights w_k > 0
    """
    return _UnrolledFn.apply(mask, meas, lambdas, tv_weights, int(tv_iter_max), bool(logits),
                             bool(measure))


class UnrolledGAPTV(torch.nn.Module):
    """Learnable unrolled GAP-TV: mask logits (or a fixed mask), lambda_k and
    tv_weight_k as parameters.  ``forward(truth)`` measures y = A(truth) with
    the current mask and reconstructs from it."""

    def __init__(self, H, W, T, iters=10, tv_iter_max=5, lambda_init=1.0, tv_weight_init=0.1,
                 mask=None, learn_mask=True, logits_init=None, device="cuda"):
        super().__init__()
        self.shape = (T, H, W)
        self.tv_iter_max = tv_iter_max
        self.learn_mask = learn_mask
        if learn_mask:
            z = logits_init if logits_init is not None else torch.randn(T, H, W)
            self.mask_logits = torch.nn.Parameter(z.to(device, torch.float32).contiguous())
        else:
            if mask is None:
                raise ValueError("a fixed mask is needed when learn_mask=False")
            self.register_buffer("phi", mask.to(device, torch.float32).contiguous())
        self.lambdas = torch.nn.Parameter(torch.full((iters,), float(lambda_init), device=device))
        self.tv_weights = torch.nn.Parameter(torch.full((iters,), float(tv_weight_init),
                                                        device=device))

    def forward(self, truth):
        m = self.mask_logits if self.learn_mask else self.phi
        return unrolled_gap_tv(m, truth, self.lambdas, self.tv_weights, self.tv_iter_max,
                               logits=self.learn_mask, measure=True)
