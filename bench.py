#!/usr/bin/env python3
"""Benchmark the fused GAP-TV reconstruction on one MI355X.

Headline workload: accelerated GAP-TV on 2048x2048x20, shifting Bernoulli(0.5)
mask, Gaussian measurement noise sigma = 0.01*max(y), lambda=1, 60 outer
iterations, tv_weight=0.05, 5 TV iterations per outer iteration, synthetic
seeded data.  One step = one full reconstruction (60 outer iterations) with
all inputs already resident in HBM.

Prints one JSON line: metric/value in pixel*frame*iter/s (H*W*T*K / time);
the roofline of the dominant kernel (the fused outer iteration) from HIP-event
timings taken in this run, with its HBM traffic from two rocprofv3 --pmc
passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a plain 1 GiB
elementwise copy; the batched mixed-size queue (64 solves) against the same
solves run one at a time; the unrolled differentiable GAP-TV (forward +
backward time, peak device memory); and a CPU leg that times the reference
algorithm's NumPy restatement on this host and holds every FP64 check.
"""

import argparse
import json
import math
import os
import re
import shutil
import sqlite3
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sci-algorithms_amd"))

METRIC = ("GAP-TV pixel·frame·iters/s + HBM GB/s; batched/unrolled vs plain at same PSNR")
UNIT = "pixel*frame*iter/s"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (H, W, T, K, tv_weight, mask, noise_rel)
    "2048x2048x20": (2048, 2048, 20, 60, 0.05, "shift", 0.01),
    "256x256x8": (256, 256, 8, 40, 0.1, "bernoulli", 0.0),
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def gap_bytes_per_pixel(T):
    """Algorithmic HBM bytes of one accelerated GAP-TV outer iteration per
    pixel: read x, Phi (T planes each), y, y1, Phi_sum; write x (T), y1."""
    return (3 * T + 4) * 4


# ---------------------------------------------------------------------------
# Headline: one full-size solve
# ---------------------------------------------------------------------------

class Solver:
    """Pre-allocated single-solve state on the device."""

    def __init__(self, y, phi, ps, iters, tv_weight, tile_rows=0):
        import torch
        from pnp_sci import _lib
        self._lib = _lib
        self.L = _lib.lib()
        T, H, W = phi.shape
        self.y, self.phi, self.ps = y, phi, ps
        p = _lib.Params()
        self.L.scigap_default_params(p, H, W, T)
        p.iters = iters
        p.tv_weight = tv_weight
        p.tv_iter_max = 5
        p.tile_rows = tile_rows
        self.p = p
        self.nbytes = self.L.scigap_solve_workspace_bytes(p)
        self.ws = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        self.x = torch.empty((T, H, W), dtype=torch.float32, device="cuda")
        self.stats = torch.empty(max(iters, 1), dtype=torch.int32, device="cuda")
        self.stream = _lib.stream_ptr()

    def __call__(self):
        self._lib.check(self.L.scigap_solve(self.p, self.y.data_ptr(), self.phi.data_ptr(),
                                            self.ps.data_ptr(), None, self.x.data_ptr(), None,
                                            None, self.stats.data_ptr(), self.ws.data_ptr(),
                                            self.nbytes, self.stream))


def stream_copy_ceiling(nbytes=1 << 30, reps=5):
    """Measured HBM ceiling of this device: the best of `reps` 1 GiB
    device-to-device copies (1 GiB read + 1 GiB written), HIP events."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / (best * 1e-3) / 1e9


def headline(args, prob):
    import numpy as np
    import torch
    from pnp_sci import _lib
    from pnp_sci.device import to_device
    H, W, T, K, tvw, _, _ = CONFIGS[args.config]
    solve = Solver(to_device(prob.y), to_device(prob.mask_thw), to_device(prob.phi_sum), K, tvw)
    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps

    # Dominant kernel, timed with HIP events on the launch stream.
    prof = _lib.kernel_profile(solve)
    n_it, ms_it = prof.get("fused_iteration", (0, 0.0))
    if n_it == 0:
        raise RuntimeError("no fused_iteration launches were timed: %s" % prof)
    bpp = gap_bytes_per_pixel(T)
    avg_it = ms_it / n_it                      # one launch covers all H*W pixels
    achieved = bpp * H * W / (avg_it * 1e-3) / 1e9
    iter_ms = sum(v[1] for k, v in prof.items() if k.startswith("fused")) / n_it
    iter_gbs = bpp * H * W / (iter_ms * 1e-3) / 1e9
    x = solve.x.cpu().numpy()
    mse = float(np.mean((x.astype(np.float64) - prob.truth_thw) ** 2))
    copy_gbs = stream_copy_ceiling()
    res = {
        "value": H * W * T * K / dt, "ms_per_step": dt * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "stream_copy_GBps": copy_gbs,
                     "frac_of_stream_copy": achieved / copy_gbs,
                     "kernel": "k_fused (one outer iteration, all tiles)",
                     "avg_launch_ms": avg_it,
                     "algorithmic_bytes_per_launch": bpp * H * W,
                     "algorithmic_bytes_per_pixel": bpp},
        "outer_iteration": {"ms": iter_ms, "algorithmic_GBps": iter_gbs,
                            "frac": iter_gbs / HBM_PEAK_GBS,
                            "kernels": {k: {"launches": v[0], "avg_ms": v[1] / max(v[0], 1)}
                                        for k, v in prof.items()}},
        "quality": {"psnr_db": 10 * math.log10(1.0 / mse),
                    "early_exit_frames": int(solve.stats[:K].sum().item())},
    }
    return res, solve


# ---------------------------------------------------------------------------
# HBM traffic of the fused kernel from rocprofv3 PMC counters
# ---------------------------------------------------------------------------

def traffic_child():
    """Run under rocprofv3: a 1 GiB elementwise copy (calibration), then three
    fused outer iterations on a 2048x2048x20 problem."""
    import torch
    from pnp_sci.device import to_device
    from pnp_sci.synthetic import make_problem
    src = torch.rand(1 << 28, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    prob = make_problem(2048, 2048, 20, seed=0, mask="shift", noise_rel=0.01)
    solve = Solver(to_device(prob.y), to_device(prob.mask_thw), to_device(prob.phi_sum), 3, 0.05)
    solve()
    torch.cuda.synchronize()


def _pmc_pass(counter, outdir):
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    cmd = [exe, "--pmc", counter, "-d", outdir, "-o", "run", "--", sys.executable,
           os.path.abspath(__file__), "--traffic-child"]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=180)
    dbs = [os.path.join(d, f) for d, _, fs in os.walk(outdir) for f in fs
           if f.endswith("_results.db")]
    if not dbs:
        raise RuntimeError("rocprofv3 --pmc %s wrote no database (rc %d): %s"
                           % (counter, r.returncode, r.stdout[-400:].decode(errors="replace")))
    rows = []
    for db in dbs:
        con = sqlite3.connect(db)
        rows += con.execute("select kernel_name, value from counters_collection "
                            "where counter_name = ?", (counter,)).fetchall()
        con.close()
    fused = [v for n, v in rows if re.search(r"k_fused<", n)]
    # torch's elementwise neg kernel; failing a name match, the largest other
    # kernel (the 1 GiB torch.rand fill writes as much as the copy does)
    others = [(v, n) for n, v in rows if "scigap" not in n]
    copy = [v for v, n in others if "neg" in n] or [max(others)[0]] if others else []
    if not fused or not copy:
        raise RuntimeError("counter rows for the fused kernel or the copy kernel are missing")
    return sum(fused) / len(fused), max(copy)


def measure_traffic():
    """HBM bytes per fused launch: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by 1 GiB over what the counter reads for the
    1 GiB copy."""
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            log("rocprofv3 --pmc %s" % counter)
            fused_kib, copy_kib = _pmc_pass(counter, os.path.join(tmp, counter))
            scale = (1 << 30) / (copy_kib * 1024.0)
            out[counter] = {"fused_KiB": fused_kib, "copy_KiB_for_1GiB": copy_kib,
                            "scale": scale, "bytes": fused_kib * 1024.0 * scale}
    out["bytes"] = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return out


# ---------------------------------------------------------------------------
# Batched mixed-size queue (64 solves) against one-at-a-time solves
# ---------------------------------------------------------------------------

def _mixed_queue(args):
    """The mixed-size queue (sizes drawn from {256x256x8, 512x512x10, 1024x1024x20,
    512x768x16}, Bernoulli(p~U(0.3,0.7)) masks) on the device."""
    from pnp_sci.device import to_device
    from pnp_sci.synthetic import mixed_queue_problem, mixed_queue_sizes
    log("generating the %d-solve mixed-size queue" % args.batch_count)
    draws = mixed_queue_sizes(seed=0, count=args.batch_count)
    probs = [mixed_queue_problem(k, size, p, seed=0) for k, (size, p) in enumerate(draws)]
    dev = [(to_device(pr.y), to_device(pr.mask_thw), to_device(pr.phi_sum), None)
           for pr in probs]
    bytes_it = sum(gap_bytes_per_pixel(pr.mask_thw.shape[0]) * pr.y.size for pr in probs)
    return draws, probs, dev, bytes_it


def _time_plan(plan, reps=3):
    import torch
    plan.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.run()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def batched_queue_only(args):
    """The batched queue alone (no comparators), for profiling its kernels."""
    from pnp_sci.batch import BatchPlan
    K, tvw = 40, 0.1
    _draws, _probs, dev, bytes_it = _mixed_queue(args)
    plan = BatchPlan(dev, iters=K, tv_weight=tvw, use_graph=True)
    t = _time_plan(plan)
    plan.close()
    return {"solves": len(dev), "outer_iters": K, "ms": t * 1e3,
            "algorithmic_GBps": bytes_it * K / t / 1e9,
            "mode": "persistent tile queue, hipGraph replay"}


def batched(args):
    import numpy as np
    import torch
    from pnp_sci.batch import BatchPlan
    K, tvw = 40, 0.1
    draws, probs, dev, bytes_it = _mixed_queue(args)
    pfi = sum(pr.mask_thw.size for pr in probs) * K       # pixel*frame*iter of the queue

    singles = [Solver(y, phi, ps, K, tvw) for (y, phi, ps, _) in dev]
    for s in singles:
        s()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in singles:
        s()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0

    # The plain GPU form: the reference loop (pnp_sci_algo.py:625-650) as
    # separate device operations, one solve at a time -- A, At and the
    # residual update as torch tensor operations, the TV as one library TV
    # call per outer iteration (no carried residual, nothing fused).
    from pnp_sci.tv import tv_device

    def plain(y, phi, ps):
        x = y.unsqueeze(0) * phi
        y1 = torch.zeros_like(y)
        for _ in range(K):
            yb = (x * phi).sum(0)
            y1 = y1 + (y - yb)
            x = x + phi * ((y1 - yb) / ps).unsqueeze(0)
            x = tv_device(x, tvw, 2e-4, 5)
        return x

    plain(*dev[0][:3])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs_plain = [plain(y, phi, ps) for (y, phi, ps, _) in dev]
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0

    # The same loop with the TV in PyTorch operations as well: a direct GPU
    # port of the reference (denoise_tv_chambolle of scikit-image 0.18.3,
    # _denoise.py:341-393, per frame with its energy stop).
    def tv_torch(img, weight, eps, n_iter_max):
        T_ = img.shape[0]
        p = torch.zeros((2,) + img.shape, device=img.device)
        g = torch.zeros_like(p)
        d = torch.zeros_like(img)
        result = torch.empty_like(img)
        done = torch.zeros(T_, dtype=torch.bool, device=img.device)
        e_init = e_prev = None
        size = img.shape[1] * img.shape[2]
        for i in range(n_iter_max):
            if i > 0:
                d = -p[0] - p[1]
                d[:, 1:] += p[0, :, :-1]
                d[:, :, 1:] += p[1, :, :, :-1]
                out = img + d
            else:
                out = img
            e = (d * d).sum((1, 2))
            g[0, :, :-1] = out[:, 1:] - out[:, :-1]
            g[1, :, :, :-1] = out[:, :, 1:] - out[:, :, :-1]
            norm = torch.sqrt((g * g).sum(0))
            e = (e + weight * norm.sum((1, 2))) / size
            p_new = (p - 0.25 * g) / (1.0 + (0.25 / weight) * norm)
            if i == 0:
                e_init = e_prev = e
            else:
                stop = (e_prev - e).abs() < eps * e_init
                newly = stop & ~done
                result = torch.where(newly[:, None, None], out, result)
                done = done | stop
                e_prev = e
            p = torch.where(done[None, :, None, None], p, p_new)
        return torch.where(done[:, None, None], result, out)

    def plain_torch(y, phi, ps):
        x = y.unsqueeze(0) * phi
        y1 = torch.zeros_like(y)
        for _ in range(K):
            yb = (x * phi).sum(0)
            y1 = y1 + (y - yb)
            x = x + phi * ((y1 - yb) / ps).unsqueeze(0)
            x = tv_torch(x, tvw, 2e-4, 5)
        return x

    plain_torch(*dev[0][:3])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs_torch = [plain_torch(y, phi, ps) for (y, phi, ps, _) in dev]
    torch.cuda.synchronize()
    t_torch = time.perf_counter() - t0

    from pnp_sci import tune
    log("autotuning the batched plan")
    best, cand_times = tune.autotune_batch(dev, iters=K, tv_weight=tvw)
    plan = BatchPlan(dev, iters=K, tv_weight=tvw, use_graph=True, tune=True)
    t_batch = _time_plan(plan)

    def mean_psnr(xs):
        v = []
        for pr, x in zip(probs, xs):
            mse = float(np.mean((x.cpu().numpy().astype(np.float64) - pr.truth_thw) ** 2))
            v.append(10 * math.log10(1.0 / mse))
        return float(np.mean(v))

    ps_seq = mean_psnr([s.x for s in singles])
    ps_batch = mean_psnr(plan.x)
    worst_vs_single = max(float(np.linalg.norm(b.cpu().numpy().astype(np.float64) -
                                               s.x.cpu().numpy()) /
                                np.linalg.norm(s.x.cpu().numpy().astype(np.float64)))
                          for b, s in zip(plan.x, singles))
    fp64 = {}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_c
        for k, (size, _p) in enumerate(draws):
            key = "%dx%dx%d" % size
            if key in fp64:
                continue
            pr = probs[k]
            want = oracle_c.gap_tv(pr.y, pr.mask, pr.phi_sum, iters=K, tv_weight=tvw)
            got = plan.x[k].cpu().numpy().astype(np.float64).transpose(1, 2, 0)
            fp64[key] = {"solve": k, "rel_err": float(np.linalg.norm(got - want) /
                                                      np.linalg.norm(want))}
    except ImportError as e:
        fp64["error"] = str(e)
    ps_plain = mean_psnr(xs_plain)
    ps_torch = mean_psnr(xs_torch)
    plan.close()
    sizes = {}
    for (size, _p) in draws:
        key = "%dx%dx%d" % size
        sizes[key] = sizes.get(key, 0) + 1
    return {
        "workload": "%d GAP-TV solves, sizes drawn from {256x256x8, 512x512x10, 1024x1024x20, "
                    "512x768x16}, Bernoulli(p~U(0.3,0.7)) masks, K=%d, tv_weight %g, 5 TV "
                    "iterations" % (len(probs), K, tvw),
        "size_counts": sizes,
        "batched": {"value": pfi / t_batch, "ms": t_batch * 1e3,
                    "algorithmic_GBps": bytes_it * K / t_batch / 1e9,
                    "mean_psnr_db": ps_batch, "mode": "persistent tile queue, hipGraph replay",
                    "tuned_options": plan.options_used,
                    "tuning_ms": {"tile_rows=%d grid=%d order=%d" % (
                        dict(k)["tile_rows"], dict(k)["grid"], dict(k)["order"]): v * 1e3
                        for k, v in cand_times.items()}},
        "one_at_a_time": {"value": pfi / t_seq, "ms": t_seq * 1e3,
                          "mean_psnr_db": ps_seq,
                          "mode": "scigap_solve per measurement, same fused kernels"},
        "plain_gpu": {"value": pfi / t_plain, "ms": t_plain * 1e3, "mean_psnr_db": ps_plain,
                      "mode": "reference loop as separate device operations per solve: "
                              "torch A/At/residual, one library TV call per iteration"},
        "plain_torch": {"value": pfi / t_torch, "ms": t_torch * 1e3, "mean_psnr_db": ps_torch,
                        "mode": "reference loop in PyTorch operations per solve, TV included "
                                "(Chambolle with the per-frame energy stop)"},
        "speedup": t_seq / t_batch, "speedup_vs_plain_gpu": t_plain / t_batch,
        "speedup_vs_plain_torch": t_torch / t_batch, "unit": UNIT,
        "speedup_bar": {"target": 3.0, "comparator": "one_at_a_time",
                        "met": t_seq / t_batch >= 3.0,
                        "note": "the same solves run one at a time through scigap_solve "
                                "(the plain GPU path of this library); plain_gpu and "
                                "plain_torch are reported beside it"},
        "batched_vs_one_at_a_time_worst_rel": worst_vs_single,
        "fp64_check": {"oracle": "FP64 C oracle, one solve of each size, K=%d" % K,
                       "per_size": fp64},
    }


# ---------------------------------------------------------------------------
# Unrolled differentiable GAP-TV
# ---------------------------------------------------------------------------

def other_solves(prob):
    """ADMM-TV on the headline problem (gamma 0.01, the reference default)
    and the small solve (256x256x8, Bernoulli(0.5), K=40, w=0.1): time
    per solve with data in HBM, and the fused kernel's algorithmic bandwidth."""
    import numpy as np
    import torch
    from pnp_sci import _lib
    from pnp_sci.device import to_device
    from pnp_sci.synthetic import problem_256x256x8
    out = {}
    cases = (("admm_2048x2048x20_K60", prob, 60, 0.05, _lib.METHOD_ADMM, (5 * 20 + 2) * 4),
             ("gap_256x256x8_K40", problem_256x256x8(0), 40, 0.1, _lib.METHOD_GAP, gap_bytes_per_pixel(8)))
    for name, pr, K, tvw, method, bpp in cases:
        solve = Solver(to_device(pr.y), to_device(pr.mask_thw), to_device(pr.phi_sum), K, tvw)
        solve.p.method = method
        solve.nbytes = solve.L.scigap_solve_workspace_bytes(solve.p)
        solve.ws = torch.empty(solve.nbytes, dtype=torch.uint8, device="cuda")
        solve()
        torch.cuda.synchronize()
        reps = 3 if pr.y.size > 1 << 20 else 20
        t0 = time.perf_counter()
        for _ in range(reps):
            solve()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        prof = _lib.kernel_profile(solve)
        n_it, ms_it = prof.get("fused_iteration", (1, float("nan")))
        T, H, W = pr.mask_thw.shape
        x = solve.x.cpu().numpy().astype(np.float64)
        rel64 = None
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_c
            want = oracle_c.gap_tv(pr.y, pr.mask, pr.phi_sum, iters=K, tv_weight=tvw,
                                   method="admm" if method == _lib.METHOD_ADMM else "gap")
            rel64 = float(np.linalg.norm(x.transpose(1, 2, 0) - want) / np.linalg.norm(want))
        except ImportError:
            pass
        out[name] = {"ms_per_solve": dt * 1e3, "pixel_frame_iter_per_s": H * W * T * K / dt,
                     "rel_err_vs_fp64_oracle": rel64,
                     "fused_launch_ms": ms_it / n_it,
                     "fused_algorithmic_GBps": bpp * H * W / (ms_it / n_it * 1e-3) / 1e9,
                     "psnr_db": 10 * math.log10(1.0 / float(np.mean((x - pr.truth_thw) ** 2)))}
    return out


def unrolled(args):
    import torch
    from pnp_sci.synthetic import unrolled_batch4_256x256x8, unrolled_batch2_1024x1024x20
    from pnp_sci.unrolled import unrolled_gap_tv
    out = {}
    for name, (data, K, logits, w0) in (("batch4_256x256x8_K10", (unrolled_batch4_256x256x8, 10, True, 0.1)),
                                        ("batch2_1024x1024x20_K20", (unrolled_batch2_1024x1024x20, 20, False, 0.05))):
        truth, mask = data(seed=0)
        tt = torch.tensor(truth, device="cuda")
        m = torch.tensor(mask, device="cuda", requires_grad=logits)
        lt = torch.full((K,), 1.0, device="cuda", requires_grad=True)
        wt = torch.full((K,), w0, device="cuda", requires_grad=True)

        x = unrolled_gap_tv(m, tt, lt, wt, logits=logits)   # warm-up
        ((x - tt) ** 2).mean().backward()
        del x
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        reps = 3
        fwd = bwd = 0.0
        for _ in range(reps):
            t0 = time.perf_counter()
            x = unrolled_gap_tv(m, tt, lt, wt, logits=logits)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ((x - tt) ** 2).mean().backward()
            torch.cuda.synchronize()
            fwd += t1 - t0
            bwd += time.perf_counter() - t1
        fwd, bwd = fwd / reps, bwd / reps
        total = fwd + bwd
        peak = torch.cuda.max_memory_allocated() - base
        # Library kernels of one more forward + backward, by kind (HIP events).
        from pnp_sci import _lib

        def step():
            xx = unrolled_gap_tv(m, tt, lt, wt, logits=logits)
            ((xx - tt) ** 2).mean().backward()
            torch.cuda.synchronize()
        prof = _lib.kernel_profile(step)
        kernels = {k: {"launches": n, "ms": ms} for k, (n, ms) in sorted(prof.items())}
        B, T, H, W = truth.shape
        out[name] = {"fwd_bwd_ms": total * 1e3, "fwd_ms": fwd * 1e3, "bwd_ms": bwd * 1e3,
                     "peak_extra_GiB": peak / 2 ** 30, "batch": B, "shape": [T, H, W],
                     "iters": K, "tv_iter_max": 5,
                     "mask": "sigmoid(logits), learnable" if logits else "shifting binary, fixed",
                     "pixel_frame_iter_per_s": B * T * H * W * K / total,
                     "kernels_one_fwd_bwd": kernels}
        del x, tt, m, lt, wt
        torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------
# CPU leg: baseline timing of the reference algorithm and the FP64 checks
# ---------------------------------------------------------------------------

def cpu_leg(args, prob, solve, want_unrolled):
    """Times the NumPy restatement of the reference path (fp32, one thread)
    on a bounded sample of the headline solve, and holds every FP64 check of
    this run: the C oracle on the full-size problem for the first
    --check-iters outer iterations, and the FP64 torch reference of the
    unrolled model (gradients of lambda_k / tv_weight_k)."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pnp_sci_oracle as O
    from pnp_sci import _lib
    from pnp_sci.device import thw_to_hwt
    H, W, T, K, tvw, _, _ = CONFIGS[args.config]
    checks = {}
    if not args.no_check:
        try:
            import oracle_c
            kc = min(args.check_iters, K) if args.check_iters > 0 else K
            log("FP64 C oracle, %d outer iterations" % kc)
            short = solve
            if kc != K:
                short = Solver(solve.y, solve.phi, solve.ps, kc, tvw)
                short()
            t1 = time.perf_counter()
            want, exits = oracle_c.gap_tv(prob.y, prob.mask, prob.phi_sum, iters=kc,
                                          tv_weight=tvw, tv_iter_max=5, exits=True)
            got = thw_to_hwt(short.x).cpu().numpy().astype(np.float64)
            checks["early_exit_frames_fp64_oracle"] = int(exits.sum())
            checks["early_exits_per_iteration_equal"] = (
                short.stats[:kc].cpu().tolist() == exits.tolist())
            checks["rel_err_vs_fp64_oracle"] = float(np.linalg.norm(got - want)
                                                     / np.linalg.norm(want))
            checks["oracle_check"] = ("full %dx%dx%d problem, %d of %d outer iterations, FP64 C "
                                      "oracle (%.0f s on the host)"
                                      % (H, W, T, kc, K, time.perf_counter() - t1))
        except ImportError as e:
            checks["rel_err_vs_fp64_oracle"] = None
            checks["oracle_check"] = str(e)
    if want_unrolled:
        import unrolled_ref as R
        from pnp_sci.synthetic import unrolled_batch2_1024x1024x20
        from pnp_sci.unrolled import unrolled_gap_tv
        log("FP64 reference gradients of the 1024x1024x20 unrolled model")
        truth, mask = unrolled_batch2_1024x1024x20(seed=0)
        Ku = 20
        tt = torch.tensor(truth, device="cuda")
        lt = torch.full((Ku,), 1.0, device="cuda", requires_grad=True)
        wt = torch.full((Ku,), 0.05, device="cuda", requires_grad=True)
        x = unrolled_gap_tv(torch.tensor(mask, device="cuda"), tt, lt, wt, logits=False)
        ((x - tt) ** 2).mean().backward()
        del x
        dd = dict(device="cuda", dtype=torch.float64)
        l64 = torch.full((Ku,), 1.0, **dd, requires_grad=True)
        w64 = torch.full((Ku,), 0.05, **dd, requires_grad=True)
        R.mse_loss(torch.tensor(mask, **dd), tt.double(), l64, w64, checkpoint=True).backward()

        def rel(a, b):
            return float((a.double() - b).norm() / b.norm())

        checks["unrolled_grad_rel_err_vs_fp64"] = {
            "lambda": rel(lt.grad, l64.grad), "tv_weight": rel(wt.grad, w64.grad),
            "reference": "plain PyTorch float64 model with per-iteration checkpoints"}
        del tt, l64, w64
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (%d outer iterations)" % args.cpu_iters)
        m = np.ascontiguousarray(prob.mask)  # the reference holds C-order (H,W,T) arrays
        t0 = time.perf_counter()
        O.gap_denoise(prob.y, prob.phi_sum, lambda v: O.A_(v, m), lambda v: O.At_(v, m),
                      _lambda=1, accelerate=True, iter_max=args.cpu_iters, tv_weight=tvw,
                      tv_iter_max=5, show_iqa=False)
        secs = time.perf_counter() - t0
        cpu = {"value": H * W * T * args.cpu_iters / secs, "unit": UNIT, "cores": 1,
               "kind": "port",
               "sample": "first %d of %d outer iterations of the same %dx%dx%d solve, NumPy "
                         "fp32 restatement of the reference path, one thread (%.1f s); host has "
                         "%d CPUs, %d in this process's affinity set"
                         % (args.cpu_iters, K, H, W, T, secs, os.cpu_count() or 0,
                            len(os.sched_getaffinity(0)))}
    return cpu, checks


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="2048x2048x20", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-iters", type=int, default=4,
                    help="outer iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 oracle comparisons")
    ap.add_argument("--check-iters", type=int, default=0,
                    help="outer iterations of the full-size FP64 oracle comparison "
                         "(0: all of them)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 PMC passes")
    ap.add_argument("--no-batch", action="store_true", help="skip the mixed-size batch")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--no-unrolled", action="store_true", help="skip the unrolled model")
    ap.add_argument("--only-unrolled", action="store_true",
                    help="run only the unrolled-model leg and print its JSON (for profiling)")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batched queue (no comparators) and print its JSON "
                         "(for profiling)")
    ap.add_argument("--only-other", action="store_true",
                    help="run only the other solves (ADMM at full size, GAP at 256x256x8) and "
                         "print their JSON (for profiling)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the GAP-TV path is built and measured on a single "
              "MI355X (sharding across GPUs is design-only); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.traffic_child:
        traffic_child()
        return
    if args.only_unrolled:
        print(json.dumps({"unrolled": unrolled(args)}), flush=True)
        return
    if args.only_batch:
        print(json.dumps({"batched_queue": batched_queue_only(args)}), flush=True)
        return

    from pnp_sci.synthetic import make_problem
    H, W, T, K, tvw, mask_kind, noise = CONFIGS[args.config]
    log("generating %s %dx%dx%d" % (args.config, H, W, T))
    prob = make_problem(H, W, T, seed=0, mask=mask_kind, noise_rel=noise)
    if args.only_other:
        print(json.dumps({"other_solves": other_solves(prob)}), flush=True)
        return
    head, solve = headline(args, prob)

    if not args.no_traffic and args.config == "2048x2048x20":
        try:
            t = measure_traffic()
            head["roofline"]["traffic"] = t["bytes"]
            head["roofline"]["traffic_detail"] = t
        except Exception as e:  # the profiler is measurement support, not the product
            head["roofline"]["traffic_error"] = repr(e)[:300]

    extra = {}
    if not args.no_batch:
        extra["batched_queue"] = batched(args)
    if not args.no_unrolled:
        extra["unrolled"] = unrolled(args)
    if args.config == "2048x2048x20" and not args.no_batch:
        extra["other_solves"] = other_solves(prob)

    cpu, checks = cpu_leg(args, prob, solve, want_unrolled=not args.no_unrolled and
                          not args.no_check)
    head["quality"].update(checks)

    out = {
        "metric": METRIC, "value": head["value"], "unit": UNIT, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "GAP-TV accelerated, %dx%dx%d, %d outer iterations, "
                               "tv_weight %g, 5 TV iterations, %s mask%s"
                               % (H, W, T, K, tvw, mask_kind,
                                  ", noise 0.01*max(y)" if noise else ""),
                   "H": H, "W": W, "T": T, "outer_iters": K, "tv_iter_max": 5},
        "roofline": head["roofline"],
        "outer_iteration": head["outer_iteration"],
        "quality": head["quality"],
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
