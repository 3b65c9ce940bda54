#!/usr/bin/env python3
"""Benchmark the fused GAP-TV reconstruction on one MI355X.

Workload (default): accelerated GAP-TV on 2048x2048x20, shifting Bernoulli(0.5)
mask, Gaussian measurement noise sigma = 0.01*max(y), lambda=1, 60 outer
iterations, tv_weight=0.05, 5 TV iterations per outer iteration, synthetic
seeded data.  One step = one full reconstruction (60 outer iterations) with
all inputs already resident in HBM.

Prints one JSON line: metric/value in pixel*frame*iter/s (H*W*T*K / time),
the roofline of the dominant kernel (fused outer iteration, interior tiles)
from HIP-event timings taken in this run, the reconstruction quality (PSNR vs
the synthetic truth, relative error vs the FP64 C oracle when it is built),
and the reference algorithm's NumPy path timed on this host as cpu_baseline.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sci-algorithms_amd"))

METRIC = ("GAP-TV pixel·frame·iters/s + HBM GB/s; batched/unrolled vs plain at same PSNR")
UNIT = "pixel*frame*iter/s"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (H, W, T, K, tv_weight, mask, noise_rel)
    "cfg2": (2048, 2048, 20, 60, 0.05, "shift", 0.01),
    "cfg1": (256, 256, 8, 40, 0.1, "bernoulli", 0.0),
}


def cpu_baseline(prob, K_sample, tv_weight):
    """The reference's NumPy path (the oracle's restatement, fp32, one thread)
    on the first K_sample outer iterations of the same solve."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pnp_sci_oracle as O
    m = np.ascontiguousarray(prob.mask)  # the reference holds C-order (H,W,T) arrays
    A = lambda v: O.A_(v, m)  # noqa: E731
    At = lambda v: O.At_(v, m)  # noqa: E731
    t0 = time.perf_counter()
    O.gap_denoise(prob.y, prob.phi_sum, A, At, _lambda=1, accelerate=True, iter_max=K_sample,
                  tv_weight=tv_weight, tv_iter_max=5, show_iqa=False)
    dt = time.perf_counter() - t0
    H, W, T = m.shape
    return H * W * T * K_sample / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="cfg2", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-iters", type=int, default=2,
                    help="outer iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 oracle comparison")
    ap.add_argument("--check-iters", type=int, default=10,
                    help="outer iterations of the full-size FP64 oracle comparison")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the GAP-TV path is built and measured on a single "
              "MI355X (sharding across GPUs is design-only); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pnp_sci import _lib
    from pnp_sci.device import hwt_to_thw, thw_to_hwt, to_device
    from pnp_sci.synthetic import make_problem

    H, W, T, K, tvw, mask_kind, noise = CONFIGS[args.config]
    print("[bench] generating %s %dx%dx%d" % (args.config, H, W, T), file=sys.stderr, flush=True)
    prob = make_problem(H, W, T, seed=0, mask=mask_kind, noise_rel=noise)

    L = _lib.lib()
    y = to_device(prob.y)
    ps = to_device(prob.phi_sum)
    phi = to_device(prob.mask_thw)
    p = _lib.Params()
    L.scigap_default_params(p, H, W, T)
    p.iters = K
    p.tv_weight = tvw
    p.tv_iter_max = 5
    nbytes = L.scigap_solve_workspace_bytes(p)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    x = torch.empty((T, H, W), dtype=torch.float32, device="cuda")
    stats = torch.empty(K, dtype=torch.int32, device="cuda")
    stream = _lib.stream_ptr()

    def solve():
        _lib.check(L.scigap_solve(p, y.data_ptr(), phi.data_ptr(), ps.data_ptr(), None,
                                  x.data_ptr(), None, None, stats.data_ptr(), ws.data_ptr(),
                                  nbytes, stream))

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    print("[bench] timing %d steps" % args.steps, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    value = H * W * T * K / dt

    # Dominant kernel, timed with HIP events on the launch stream.
    prof = _lib.kernel_profile(solve)
    n_it, ms_it = prof.get("fused_iteration", (0, 0.0))
    if n_it == 0:
        raise RuntimeError("no fused_iteration launches were timed: %s" % prof)
    bytes_px = (3 * T + 4) * 4                 # accelerated GAP, per pixel per outer iteration
    avg_it = ms_it / n_it                      # one launch covers all H*W pixels
    achieved = bytes_px * H * W / (avg_it * 1e-3) / 1e9
    iter_ms = sum(v[1] for k, v in prof.items() if k.startswith("fused")) / n_it
    iter_gbs = bytes_px * H * W / (iter_ms * 1e-3) / 1e9

    xo = x.cpu().numpy()
    mse = float(np.mean((xo.astype(np.float64) - prob.truth_thw) ** 2))
    quality = {"psnr_db": 10 * math.log10(1.0 / mse),
               "early_exit_frames": int(stats.sum().item())}
    if not args.no_check:
        # FP64 C oracle on the same full-size problem for the first check-iters
        # outer iterations (the full 60 take minutes of host time).
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_c
            kc = min(args.check_iters, K)
            print("[bench] FP64 C oracle, %d outer iterations" % kc, file=sys.stderr, flush=True)
            pc = _lib.Params.from_buffer_copy(p)
            pc.iters = kc
            xc = torch.empty_like(x)
            sc = torch.empty(kc, dtype=torch.int32, device="cuda")
            _lib.check(L.scigap_solve(pc, y.data_ptr(), phi.data_ptr(), ps.data_ptr(), None,
                                      xc.data_ptr(), None, None, sc.data_ptr(), ws.data_ptr(),
                                      nbytes, stream))
            t1 = time.perf_counter()
            want = oracle_c.gap_tv(prob.y, prob.mask, prob.phi_sum, iters=kc, tv_weight=tvw,
                                   tv_iter_max=5)
            got = thw_to_hwt(xc).cpu().numpy().astype(np.float64)
            quality["rel_err_vs_fp64_oracle"] = float(
                np.linalg.norm(got - want) / np.linalg.norm(want))
            quality["oracle_check"] = ("full %dx%dx%d problem, first %d outer iterations, "
                                       "FP64 C oracle (%.0f s on the host)"
                                       % (H, W, T, kc, time.perf_counter() - t1))
        except ImportError as e:
            quality["rel_err_vs_fp64_oracle"] = None
            quality["oracle_check"] = str(e)

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (%d outer iterations)" % args.cpu_iters, file=sys.stderr,
              flush=True)
        v, secs = cpu_baseline(prob, args.cpu_iters, tvw)
        cpu = {"value": v, "unit": UNIT, "cores": 1, "kind": "port",
               "sample": "first %d of %d outer iterations of the same %dx%dx%d solve, NumPy "
                         "fp32 restatement of the reference path, one thread (%.1f s); "
                         "host has %d CPUs" % (args.cpu_iters, K, H, W, T, secs,
                                               os.cpu_count() or 0)}

    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "GAP-TV accelerated, %dx%dx%d, %d outer iterations, "
                               "tv_weight %g, 5 TV iterations, %s mask%s"
                               % (H, W, T, K, tvw, mask_kind,
                                  ", noise 0.01*max(y)" if noise else ""),
                   "H": H, "W": W, "T": T, "outer_iters": K, "tv_iter_max": 5},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused (one outer iteration, all tiles)",
                     "avg_launch_ms": avg_it,
                     "algorithmic_bytes_per_launch": bytes_px * H * W},
        "outer_iteration": {"ms": iter_ms, "algorithmic_GBps": iter_gbs,
                            "frac": iter_gbs / HBM_PEAK_GBS,
                            "kernels": {k: {"launches": v[0], "avg_ms": v[1] / max(v[0], 1)}
                                        for k, v in prof.items()}},
        "quality": quality,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
