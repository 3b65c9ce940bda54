"""Batched GAP-TV / ADMM-TV over many measurements of uneven sizes.

The reference solves coded frames and measurement blocks one after another
(PnP_SCI/python/pnp_sci_algo.py:498, pnp_sci_test_data_realexp_blkproc.py:93).
``BatchPlan`` runs a whole list of independent solves with one persistent
kernel launch per outer iteration, optionally replayed as a hipGraph
(scigap_batch_plan_create / scigap_batch_run).
"""

from __future__ import annotations

import ctypes

from . import _lib
from .device import thw_to_hwt, to_device, to_device_thw, to_host_like

__all__ = ["BatchPlan", "gap_denoise_batch"]


def _params(H, W, T, *, method, iters, _lambda, accelerate, gamma, clip, tv_weight, tv_iter_max,
            tv_eps):
    L = _lib.lib()
    p = _lib.Params()
    L.scigap_default_params(p, H, W, T)
    p.method = _lib.METHOD_ADMM if method == "admm" else _lib.METHOD_GAP
    p.accelerate = 1 if accelerate else 0
    p.clip = 1 if clip else 0
    p.iters = int(iters)
    p.lambda_ = float(_lambda)
    p.gamma = float(gamma)
    p.tv_weight = float(tv_weight)
    p.tv_iter_max = int(tv_iter_max)
    p.tv_eps = float(tv_eps)
    return p


class BatchPlan:
    """A reusable batch of solves on device tensors in library layout.

    ``problems``: list of (y (H,W), phi (T,H,W), phi_sum (H,W), x0 (T,H,W) or
    None) float32 CUDA tensors.  Outputs are allocated here (``self.x``).
    Solver options are shared by all solves.

    ``tune=True`` times the plan options on this batch first
    (``pnp_sci.tune.autotune_batch``) and uses the fastest; with the default
    options (tile_rows = grid = order = 0) a choice remembered for the same
    batch signature is used without tuning again.  ``self.options_used``
    holds the options of the plan.
    """

    def __init__(self, problems, *, method="gap", iters=50, _lambda=1.0, accelerate=True,
                 gamma=0.01, clip=False, tv_weight=0.1, tv_iter_max=5, tv_eps=2e-4,
                 use_graph=True, tile_rows=0, grid=0, order=0, tune=False, stream=None):
        import torch
        L = _lib.lib()
        if tune or (tile_rows == 0 and grid == 0 and order == 0):
            from . import tune as _tune
            solver = dict(method=method, iters=iters, _lambda=_lambda, accelerate=accelerate,
                          gamma=gamma, clip=clip, tv_weight=tv_weight, tv_iter_max=tv_iter_max,
                          tv_eps=tv_eps)
            best = _tune.tuned_batch_options(problems, solver)
            if best is None and tune:
                best, _ = _tune.autotune_batch(problems, **solver)
            if best is not None:
                tile_rows, grid, order = best["tile_rows"], best["grid"], best["order"]
        self.options_used = dict(tile_rows=tile_rows, grid=grid, order=order)
        self._L = L
        self._keep = []
        self.x = []
        n = len(problems)
        if n == 0:
            raise ValueError("empty batch")
        items = (_lib.BatchItem * n)()
        for i, (y, phi, ps, x0) in enumerate(problems):
            T, H, W = phi.shape
            out = torch.empty((T, H, W), dtype=torch.float32, device=phi.device)
            self.x.append(out)
            self._keep += [y, phi, ps, x0]
            items[i] = _lib.BatchItem(H, W, T, y.data_ptr(), phi.data_ptr(), ps.data_ptr(),
                                      x0.data_ptr() if x0 is not None else None, out.data_ptr())
        T0, H0, W0 = problems[0][1].shape
        self.params = _params(H0, W0, T0, method=method, iters=iters, _lambda=_lambda,
                              accelerate=accelerate, gamma=gamma, clip=clip, tv_weight=tv_weight,
                              tv_iter_max=tv_iter_max, tv_eps=tv_eps)
        self.options = _lib.BatchOptions(1 if use_graph else 0, int(tile_rows), int(grid),
                                         int(order))
        nbytes = L.scigap_batch_workspace_bytes(self.params, items, n, self.options)
        if nbytes == 0:
            raise _lib.ScigapError(-1, L.scigap_last_error().decode(errors="replace"))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=problems[0][1].device)
        self._items = items
        self._plan = ctypes.c_void_p()
        _lib.check(L.scigap_batch_plan_create(ctypes.byref(self._plan), self.params, items, n,
                                              self.ws.data_ptr(), nbytes, self.options,
                                              _lib.stream_ptr(stream)))

    def run(self, stream=None):
        _lib.check(self._L.scigap_batch_run(self._plan, _lib.stream_ptr(stream)))
        return self.x

    def close(self):
        if self._plan:
            self._L.scigap_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


def gap_denoise_batch(ys, masks, phi_sums, *, x0s=None, method="gap", iter_max=50, _lambda=1,
                      accelerate=True, gamma=0.01, clip=False, tv_weight=0.1, tv_iter_max=5,
                      use_graph=True):
    """Reconstruct a list of measurements (each (H,W) with an (H,W,T) mask, the
    reference's layout) in one batch; returns the list of (H,W,T) results in
    the container kind of the inputs."""
    probs = []
    for i, (y, m, ps) in enumerate(zip(ys, masks, phi_sums)):
        x0 = x0s[i] if x0s is not None and x0s[i] is not None else None
        probs.append((to_device(y), to_device_thw(m), to_device(ps),
                      to_device_thw(x0) if x0 is not None else None))
    plan = BatchPlan(probs, method=method, iters=iter_max, _lambda=_lambda, accelerate=accelerate,
                     gamma=gamma, clip=clip, tv_weight=tv_weight, tv_iter_max=tv_iter_max,
                     use_graph=use_graph)
    xs = plan.run()
    out = [to_host_like(thw_to_hwt(x), y) for x, y in zip(xs, ys)]
    plan.close()
    return out
