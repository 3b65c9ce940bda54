"""Batched solves (mixed-size queue): against single solves and the FP64 oracle.

Every result of a batch must match the same problem solved alone through
scigap_solve (same kernels; only the order of the fp64 energy atomics can
differ) to 1e-6 relative, and the FP64 oracle to 1e-4.  hipGraph replays
are each checked against an eager run, with eager work in between.
"""

import numpy as np
import pytest

import pnp_sci_oracle as O

pytestmark = pytest.mark.gpu

SIZES = [(48, 80, 5), (64, 64, 8), (100, 150, 6), (33, 47, 4), (130, 260, 8), (20, 24, 3)]


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _problems(seed=0):
    from pnp_sci.synthetic import make_problem
    return [make_problem(H, W, T, seed=seed + i, p=0.3 + 0.08 * i) for i, (H, W, T) in
            enumerate(SIZES)]


def _device(pr):
    from pnp_sci.device import to_device
    return (to_device(pr.y), to_device(pr.mask_thw), to_device(pr.phi_sum), None)


@pytest.mark.parametrize("method,kw", [("gap", {}), ("gap", {"accelerate": False}),
                                       ("admm", {}), ("admm", {"gamma": 0.0, "clip": True}),
                                       ("gap", {"tv_iter_max": 3}), ("gap", {"tv_iter_max": 7})])
def test_batch_matches_single_and_oracle(gpu, method, kw):
    from pnp_sci.batch import BatchPlan
    from pnp_sci.pnp_sci_algo import solve_device
    probs = _problems()
    dev = [_device(p) for p in probs]
    iters = 10
    plan = BatchPlan(dev, method=method, iters=iters, tv_weight=0.1, use_graph=False, **kw)
    xs = [x.cpu().numpy() for x in plan.run()]
    plan.close()
    for pr, d, xb in zip(probs, dev, xs):
        xs1, _, _ = solve_device(d[0], d[1], d[2], method=method, iters=iters, tv_weight=0.1, **kw)
        assert rel(xb, xs1.cpu().numpy()) <= 1e-6
        m = pr.mask.astype(np.float64)
        A = lambda v: O.A_(v, m)  # noqa: E731
        At = lambda v: O.At_(v, m)  # noqa: E731
        okw = dict(kw)
        if method == "gap":
            want = O.gap_denoise(pr.y.astype(np.float64), pr.phi_sum.astype(np.float64), A, At,
                                 iter_max=iters, tv_weight=0.1, show_iqa=False, **okw)[0]
        else:
            want = O.admm_denoise(pr.y.astype(np.float64), pr.phi_sum.astype(np.float64), A, At,
                                  iter_max=iters, tv_weight=0.1, show_iqa=False, **okw)[0]
        assert rel(np.moveaxis(xb, 0, -1), want) <= 1e-4


@pytest.mark.parametrize("method", ["gap", "admm"])
@pytest.mark.parametrize("n", [5, 6, 7])
def test_batch_with_many_frames(gpu, method, n):
    """A queue holding a 1030-frame solve beside an ordinary one: the batched
    init and fused kernels accumulate the frame sums in fp64 like the single
    solve -- in registers for tv_iter_max 5, as fp32 partials over blocks of 8
    frames added to fp64 LDS slots for 6 and 7 (eps = 1 makes every frame stop
    at the first exit test in FP32 and FP64 alike, so the comparison cannot
    hinge on a borderline decision)."""
    from pnp_sci.batch import BatchPlan
    from pnp_sci.synthetic import make_problem
    probs = [make_problem(9, 7, 1030, seed=6, p=0.5), make_problem(48, 80, 5, seed=1, p=0.5)]
    dev = [_device(p) for p in probs]
    iters = 4
    plan = BatchPlan(dev, method=method, iters=iters, tv_weight=0.1, tv_iter_max=n, tv_eps=1.0,
                     use_graph=False)
    xs = [x.cpu().numpy() for x in plan.run()]
    plan.close()
    for pr, xb in zip(probs, xs):
        m = pr.mask.astype(np.float64)
        run = O.gap_denoise if method == "gap" else O.admm_denoise
        want = run(pr.y.astype(np.float64), pr.phi_sum.astype(np.float64),
                   lambda v: O.A_(v, m), lambda v: O.At_(v, m), iter_max=iters,
                   tv_weight=0.1, tv_iter_max=n, show_iqa=False, _tv_eps=1.0)[0]
        assert rel(np.moveaxis(xb, 0, -1), want) <= 1e-4


@pytest.mark.parametrize("tile_rows", [8, 12, 16])
def test_batch_graph_replays_and_tile_shapes(gpu, tile_rows):
    import torch
    from pnp_sci.batch import BatchPlan
    probs = _problems(seed=7)
    dev = [_device(p) for p in probs]
    eager = BatchPlan(dev, iters=8, tv_weight=0.1, use_graph=False, tile_rows=tile_rows)
    ref = [x.clone() for x in eager.run()]
    graph = BatchPlan(dev, iters=8, tv_weight=0.1, use_graph=True, tile_rows=tile_rows)
    for rep in range(3):
        got = graph.run()
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert float((a - b).abs().max()) <= 1e-5 * float(b.abs().max()), rep
        # eager work on the same stream between replays
        again = eager.run()
        torch.cuda.synchronize()
        for a, b in zip(again, ref):
            assert float((a - b).abs().max()) <= 1e-5 * float(b.abs().max())
    eager.close()
    graph.close()


def test_gap_denoise_batch_reference_layout(gpu):
    from pnp_sci.batch import gap_denoise_batch
    probs = _problems(seed=3)[:3]
    xs = gap_denoise_batch([p.y for p in probs], [np.ascontiguousarray(p.mask) for p in probs],
                           [p.phi_sum for p in probs], iter_max=6, tv_weight=0.1)
    for pr, x in zip(probs, xs):
        m = pr.mask.astype(np.float64)
        want = O.gap_denoise(pr.y.astype(np.float64), pr.phi_sum.astype(np.float64),
                             lambda v: O.A_(v, m), lambda v: O.At_(v, m), iter_max=6,
                             tv_weight=0.1, show_iqa=False)[0]
        assert x.shape == pr.mask.shape and rel(x, want) <= 1e-4


def test_batch_rejects_unsupported(gpu):
    from pnp_sci import _lib
    from pnp_sci.batch import BatchPlan
    dev = [_device(p) for p in _problems()[:2]]
    with pytest.raises(_lib.ScigapError):
        BatchPlan(dev, iters=4, tv_iter_max=12)   # deep TV is not batched
    with pytest.raises(ValueError):
        BatchPlan([], iters=4)


def test_batch_autotune_picks_and_remembers(gpu):
    """BatchPlan(tune=True) times the plan options on the batch, uses the
    fastest, and a later plan of the same batch with default options reuses
    the choice; the tuned plan's results equal the default plan's."""
    from pnp_sci import tune
    from pnp_sci.batch import BatchPlan
    probs = _problems(seed=3)
    dev = [_device(p) for p in probs]
    tune.clear()
    base = BatchPlan(dev, iters=6, tv_weight=0.1, use_graph=False, tile_rows=8, grid=0, order=1)
    want = [x.cpu().numpy() for x in base.run()]
    base.close()
    cands = [dict(tile_rows=r, grid=g, order=o) for r in (8, 16) for g in (0, 128) for o in (0,)]
    best, times = tune.autotune_batch(dev, candidates=cands, iters=6, tv_weight=0.1)
    assert best in cands and len(times) == len(cands)
    plan = BatchPlan(dev, iters=6, tv_weight=0.1, use_graph=True, tune=True)
    assert plan.options_used == best          # remembered: no second tuning run
    got = [x.cpu().numpy() for x in plan.run()]
    plan.close()
    again = BatchPlan(dev, iters=6, tv_weight=0.1, use_graph=False)
    assert again.options_used == best
    again.close()
    for g, w in zip(got, want):
        assert rel(g, w) <= 1e-6
    tune.clear()
    assert tune.tuned_batch_options(dev, dict(method="gap", iters=6, _lambda=1.0,
                                              accelerate=True, gamma=0.01, clip=False,
                                              tv_weight=0.1, tv_iter_max=5,
                                              tv_eps=2e-4)) is None


@pytest.mark.parametrize("method", ["gap", "admm"])
def test_batch_one_iteration(gpu, method):
    """One outer iteration in a batch (ADMM then runs no fused iteration: its
    result is the projection X_0 + lambda*r_0*Phi), with and without a warm
    start, against the FP64 oracle."""
    from pnp_sci.batch import BatchPlan
    import torch
    probs = _problems(seed=5)[:3]
    dev = [_device(p) for p in probs]
    rng = np.random.default_rng(5)
    x0s = [rng.random(p.mask_thw.shape).astype(np.float32) for p in probs]
    dev[1] = (dev[1][0], dev[1][1], dev[1][2], torch.from_numpy(x0s[1]).cuda())
    plan = BatchPlan(dev, method=method, iters=1, tv_weight=0.1, use_graph=False)
    xs = [x.cpu().numpy() for x in plan.run()]
    plan.close()
    for i, (pr, xb) in enumerate(zip(probs, xs)):
        m = pr.mask.astype(np.float64)
        run = O.gap_denoise if method == "gap" else O.admm_denoise
        x0 = np.moveaxis(x0s[1], 0, -1).astype(np.float64) if i == 1 else None
        want = run(pr.y.astype(np.float64), pr.phi_sum.astype(np.float64),
                   lambda v: O.A_(v, m), lambda v: O.At_(v, m), iter_max=1, tv_weight=0.1,
                   x0=x0, show_iqa=False)[0]
        assert rel(np.moveaxis(xb, 0, -1), want) <= 1e-5, i


def test_random_batches_against_fp64_c_oracle(gpu):
    """Seeded random batches (1-8 solves of random sizes, GAP or ADMM, one to
    six iterations, TV depths 2-7, exit tolerances, tile heights, queue order,
    graph or plain launches): every solve within 1e-4 of the FP64 C oracle."""
    from pnp_sci.batch import BatchPlan
    from pnp_sci.synthetic import make_problem
    oracle_c = pytest.importorskip("oracle_c", reason="make -C oracle builds the C oracle")
    rng = np.random.default_rng(3)
    for case in range(20):
        method = ["gap", "admm"][int(rng.integers(0, 2))]
        n, K = int(rng.integers(2, 8)), int(rng.integers(1, 7))
        eps, w = float(rng.choice([2e-4, 2e-2, 1.0])), float(rng.choice([0.05, 0.1, 0.3]))
        rows = int(rng.choice([0, 8, 12, 16])) if n == 5 else 0
        graph, order = bool(rng.integers(0, 2)), int(rng.integers(0, 2))
        probs = [make_problem(int(rng.integers(1, 300)), int(rng.integers(1, 300)),
                              int(rng.integers(1, 25)), seed=100 * case + i,
                              p=float(rng.uniform(0.3, 0.7)))
                 for i in range(int(rng.integers(1, 9)))]
        plan = BatchPlan([_device(pr) for pr in probs], method=method, iters=K, tv_weight=w,
                         tv_iter_max=n, tv_eps=eps, use_graph=graph, tile_rows=rows, grid=0,
                         order=order)
        xs = [x.cpu().numpy() for x in plan.run()]
        plan.close()
        for i, (pr, xb) in enumerate(zip(probs, xs)):
            want = oracle_c.gap_tv(pr.y, pr.mask, pr.phi_sum, iters=K, method=method,
                                   tv_weight=w, eps=eps, tv_iter_max=n)
            e = rel(np.moveaxis(xb, 0, -1), want)
            assert e <= 1e-4, (case, i, pr.mask.shape, method, n, K, eps, w, rows, graph, e)
