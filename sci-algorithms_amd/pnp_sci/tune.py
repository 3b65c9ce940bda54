"""Autotuning of the fused kernels for a given problem shape.

* ``autotune_solve``: tile height x engine of the fused GAP/ADMM-TV kernel,
  timed on the caller's data by ``scigap_autotune_solve``; later solves of
  the same shape with default options use the choice.
* ``autotune_tv``: the TV fusion depth (dual updates per launch) of
  ``scigap_tv_chambolle`` for deep TV (``scigap_autotune_tv``).
* ``autotune_batch``: tile height, persistent grid and queue order of a
  batched plan, timed on the batch itself.

The C++ tuner keeps its choices in a process-wide cache, persisted to the
file named by ``SCIGAP_TUNE_CACHE`` when that variable is set.  No choice
changes a result; the tests check that every candidate agrees.
"""

from __future__ import annotations

import ctypes
import time

from . import _lib

__all__ = ["autotune_solve", "autotune_tv", "autotune_batch", "clear"]


def _result(r):
    return {"tile_rows": r.tile_rows, "engine": r.engine, "tv_depth": r.tv_depth,
            "candidates": r.candidates, "best_ms": r.best_ms, "default_ms": r.default_ms}


def autotune_solve(y, phi, phi_sum, *, method="gap", accelerate=True, tv_iter_max=5,
                   tv_weight=0.1, iters=4):
    """Tune the fused kernel on device tensors y (H,W), phi (T,H,W),
    phi_sum (H,W); returns the choice and its time per outer iteration."""
    import torch
    L = _lib.lib()
    T, H, W = phi.shape
    p = _lib.Params()
    L.scigap_default_params(p, H, W, T)
    p.method = _lib.METHOD_ADMM if method == "admm" else _lib.METHOD_GAP
    p.accelerate = 1 if accelerate else 0
    p.tv_iter_max = int(tv_iter_max)
    p.tv_weight = float(tv_weight)
    p.iters = int(iters)
    nbytes = L.scigap_solve_workspace_bytes(p)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=phi.device)
    x = torch.empty_like(phi)
    r = _lib.TuneResult()
    _lib.check(L.scigap_autotune_solve(p, y.data_ptr(), phi.data_ptr(), phi_sum.data_ptr(),
                                       x.data_ptr(), int(iters), ws.data_ptr(), nbytes,
                                       _lib.stream_ptr(), ctypes.byref(r)))
    return _result(r)


def autotune_tv(image_thw, *, weight=0.1, eps=2e-4, n_iter_max=30):
    """Tune the TV fusion depth on a (T,H,W) device tensor."""
    import torch
    L = _lib.lib()
    T, H, W = image_thw.shape
    nbytes = L.scigap_tv_workspace_bytes_depth(H, W, T, int(n_iter_max), 1)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=image_thw.device)
    out = torch.empty_like(image_thw)
    r = _lib.TuneResult()
    _lib.check(L.scigap_autotune_tv(image_thw.data_ptr(), out.data_ptr(), H, W, T, float(weight),
                                    float(eps), int(n_iter_max), ws.data_ptr(), nbytes,
                                    _lib.stream_ptr(), ctypes.byref(r)))
    return _result(r)


# Batch choices of this process: {batch signature: options}.
_batch_cache = {}


def batch_signature(problems, solver):
    """Key of a tuned batch: the shapes in submission order and the solver
    options that change the kernels' work (method, TV depth, iterations)."""
    shapes = tuple(tuple(int(v) for v in p[1].shape) for p in problems)
    # BatchPlan's defaults for options the caller left out
    defaults = (("method", "gap"), ("accelerate", True), ("tv_iter_max", 5), ("iters", 50))
    return shapes, tuple((k, solver.get(k, d)) for k, d in defaults)


def tuned_batch_options(problems, solver):
    """Options remembered for this batch by autotune_batch, or None."""
    return _batch_cache.get(batch_signature(problems, solver))


def autotune_batch(problems, *, reps=2, candidates=None, **solver):
    """Time BatchPlan options (tile height of the fused kernel, persistent
    grid, queue order) on the batch itself, remember the fastest for this
    batch signature (used by BatchPlan(..., tune=True) and by later plans of
    the same batch with default options), and return (best options,
    {options: seconds per run}).  ``problems`` as for BatchPlan.  The choice
    never changes a result (the same tile body runs in every candidate)."""
    import torch
    from .batch import BatchPlan
    if candidates is None:
        candidates = [dict(tile_rows=r, grid=g, order=o)
                      for r in (8, 12, 16) for g in (0, 256) for o in (0, 1)]
    times = {}
    for c in candidates:
        plan = BatchPlan(problems, use_graph=True, tune=False, **c, **solver)
        plan.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.run()
        torch.cuda.synchronize()
        times[tuple(sorted(c.items()))] = (time.perf_counter() - t0) / reps
        plan.close()
    best = min(times, key=times.get)
    _batch_cache[batch_signature(problems, solver)] = dict(best)
    return dict(best), times


def clear():
    """Forget every tuned choice of this process."""
    _batch_cache.clear()
    _lib.check(_lib.lib().scigap_tune_clear())
