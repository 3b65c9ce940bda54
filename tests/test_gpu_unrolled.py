"""Differentiable unrolled GAP-TV: hand-written HIP backward against the
plain-PyTorch FP64 reference (oracle/unrolled_ref.py, pinned on the CPU to the
NumPy oracle) and against FP64 central finite differences.

Tolerances: the forward pass is fp32 against fp64, within 1e-4 relative;
gradients within 1e-3 relative L2 of the FP64 autograd gradients, and the
32 directional derivatives of the batch-4 256x256x8 problem within 1e-3
(relative L2 over the directions) of FP64 central differences with eps=1e-6.
The reference runs in float64 on the same device as a checker.
"""

import numpy as np
import pytest

import unrolled_ref as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _schedules(K, seed):
    g = np.random.default_rng(seed)
    lam = (1.0 + 0.2 * g.standard_normal(K)).astype(np.float32)
    w = (0.05 + 0.1 * g.random(K)).astype(np.float32)
    return lam, w


def _hip(mask, truth, lam, w, n, logits, measure=True, y=None):
    """x, loss and gradients from the HIP path (MSE loss against truth)."""
    import torch
    from pnp_sci.unrolled import unrolled_gap_tv
    m = torch.tensor(mask, device="cuda", requires_grad=True)
    tt = torch.tensor(truth, device="cuda")
    lt = torch.tensor(lam, device="cuda", requires_grad=True)
    wt = torch.tensor(w, device="cuda", requires_grad=True)
    meas = tt if measure else torch.tensor(y, device="cuda", requires_grad=True)
    x = unrolled_gap_tv(m, meas, lt, wt, tv_iter_max=n, logits=logits, measure=measure)
    loss = ((x - tt) ** 2).mean()
    loss.backward()
    out = {"x": x.detach().cpu().numpy(), "loss": loss.item(), "mask": m.grad.cpu().numpy(),
           "lam": lt.grad.cpu().numpy(), "w": wt.grad.cpu().numpy()}
    if not measure:
        out["y"] = meas.grad.cpu().numpy()
    return out


def _ref(mask, truth, lam, w, n, logits, measure=True, y=None, checkpoint=False):
    """The same quantities from the FP64 torch reference (on the device)."""
    import torch
    dd = dict(device="cuda", dtype=torch.float64)
    m = torch.tensor(mask, **dd, requires_grad=True)
    tt = torch.tensor(truth, **dd)
    lt = torch.tensor(lam, **dd, requires_grad=True)
    wt = torch.tensor(w, **dd, requires_grad=True)
    phi = torch.sigmoid(m) if logits else m
    yy = R.measure(phi, tt) if measure else torch.tensor(y, **dd, requires_grad=True)
    x = R.unrolled_gap_tv(phi, yy, lt, wt, tv_iter_max=n, checkpoint=checkpoint)
    loss = ((x - tt) ** 2).mean()
    loss.backward()
    out = {"x": x.detach().cpu().numpy(), "loss": loss.item(), "mask": m.grad.cpu().numpy(),
           "lam": lt.grad.cpu().numpy(), "w": wt.grad.cpu().numpy()}
    if not measure:
        out["y"] = yy.grad.cpu().numpy()
    return out


CASES = [
    # (B, H, W, T, K, n, logits)
    (1, 37, 50, 3, 3, 5, True),      # ragged: border tiles only, odd sizes
    (2, 64, 96, 4, 4, 5, True),
    (2, 130, 260, 6, 3, 3, True),    # several tiles, depth 2
    (1, 48, 64, 5, 2, 7, False),     # fixed mask, depth 6
    (3, 40, 40, 2, 5, 2, False),     # depth 1
    (1, 150, 270, 3, 2, 6, True),    # depth 5, interior and border backward tiles
    (2, 100, 200, 3, 2, 4, False),   # depth 3
    (2, 5, 9, 2, 2, 5, True),        # frames smaller than the reverse tile's halo
    (1, 9, 7, 600, 2, 5, False),     # 600 frames: the frame sums of forward and reverse
    (1, 300, 500, 2, 2, 5, False),   # edge-aligned reverse tiles, middle tiles both ways
    (1, 201, 388, 3, 2, 5, True),    # odd height: rows unaligned, columns aligned
]


@pytest.mark.parametrize("B,H,W,T,K,n,logits", CASES)
def test_forward_and_gradients_match_fp64_reference(gpu, B, H, W, T, K, n, logits):
    from pnp_sci.synthetic import unrolled_batch
    truth, mask = unrolled_batch(B, H, W, T, seed=H + K, mask="logits" if logits else "shift")
    if not logits:
        mask = 0.2 + 0.6 * mask   # a non-binary fixed mask also exercises dL/dphi
    lam, w = _schedules(K, H)
    got = _hip(mask, truth, lam, w, n, logits)
    want = _ref(mask, truth, lam, w, n, logits)
    print("case", (B, H, W, T, K, n, logits), {k: rel(got[k], want[k]) for k in got})
    assert rel(got["x"], want["x"]) <= 1e-4
    assert abs(got["loss"] - want["loss"]) <= 1e-4 * want["loss"]
    for k in ("mask", "lam", "w"):
        assert rel(got[k], want[k]) <= 1e-3, k


def test_measurement_gradient(gpu):
    """measure=0: the measurement y is an input and receives dL/dy."""
    from pnp_sci.synthetic import unrolled_batch
    truth, z = unrolled_batch(2, 56, 72, 4, seed=11)
    phi = 1.0 / (1.0 + np.exp(-z.astype(np.float64)))
    y = np.einsum("thw,bthw->bhw", phi, truth).astype(np.float32)
    lam, w = _schedules(3, 11)
    got = _hip(phi.astype(np.float32), truth, lam, w, 5, logits=False, measure=False, y=y)
    want = _ref(phi, truth, lam, w, 5, logits=False, measure=False, y=y)
    for k in ("x", "y", "mask", "lam", "w"):
        assert rel(got[k], want[k]) <= 1e-3, (k, rel(got[k], want[k]))


def test_unrolled_batch4_256x256x8_finite_differences(gpu):
    """Batch 4 at 256x256x8, K=10 x 5 TV iterations, mask logits ~ N(0,1):
    the hand-written gradient against FP64 central differences (eps=1e-6) of
    the FP64 reference loss on 32 random unit directions in (logits,
    lambda_k, tv_weight_k), and against FP64 autograd."""
    import torch
    from pnp_sci.synthetic import unrolled_batch4_256x256x8
    truth, z = unrolled_batch4_256x256x8(seed=0)
    K, n = 10, 5
    lam, w = _schedules(K, 0)
    got = _hip(z, truth, lam, w, n, logits=True)
    want = _ref(z, truth, lam, w, n, logits=True)
    errs = {k: rel(got[k], want[k]) for k in ("x", "mask", "lam", "w")}
    print("unrolled_batch4_256x256x8 relative errors vs FP64 autograd", errs)
    for k, e in errs.items():
        assert e <= (1e-4 if k == "x" else 1e-3), (k, e)

    dd = dict(device="cuda", dtype=torch.float64)
    tt = torch.tensor(truth, **dd)
    z0 = torch.tensor(z, **dd)
    l0 = torch.tensor(lam, **dd)
    w0 = torch.tensor(w, **dd)

    def loss(dz, dl, dw, s):
        with torch.no_grad():
            return R.mse_loss(torch.sigmoid(z0 + s * dz), tt, l0 + s * dl, w0 + s * dw,
                              tv_iter_max=n).item()

    g = torch.Generator(device="cpu").manual_seed(1234)
    eps = 1e-6
    fds, dirs = [], []
    for _ in range(32):
        # a unit direction with equal weight on each parameter block
        blocks = [torch.randn(z.shape, generator=g, dtype=torch.float64),
                  torch.randn(K, generator=g, dtype=torch.float64),
                  torch.randn(K, generator=g, dtype=torch.float64)]
        blocks = [b / b.norm() / np.sqrt(3.0) for b in blocks]
        dz, dl, dw = (b.to("cuda") for b in blocks)
        fds.append((loss(dz, dl, dw, eps) - loss(dz, dl, dw, -eps)) / (2 * eps))
        dirs.append(float((torch.from_numpy(got["mask"]).double() * blocks[0]).sum()
                          + (torch.from_numpy(got["lam"]).double() * blocks[1]).sum()
                          + (torch.from_numpy(got["w"]).double() * blocks[2]).sum()))
    fds, dirs = np.array(fds), np.array(dirs)
    err = rel(dirs, fds)
    worst = float(np.max(np.abs(dirs - fds) / np.abs(fds)))
    print("unrolled_batch4_256x256x8 FD check: relative L2 %.3e, worst direction %.3e" % (err, worst))
    assert err <= 1e-3


def test_unrolled_batch2_1024x1024x20_checkpointed_gradients(gpu):
    """1024x1024x20, K=20 x 5, batch 2, shifting binary mask, learnable
    lambda_k / tv_weight_k: gradients against the FP64 reference run with
    per-outer-iteration checkpointing."""
    import torch
    from pnp_sci.synthetic import unrolled_batch2_1024x1024x20
    from pnp_sci.unrolled import unrolled_gap_tv
    truth, mask = unrolled_batch2_1024x1024x20(seed=0)
    K = 20
    lam, w = np.full(K, 1.0, np.float32), np.full(K, 0.05, np.float32)
    tt = torch.tensor(truth, device="cuda")
    lt = torch.tensor(lam, device="cuda", requires_grad=True)
    wt = torch.tensor(w, device="cuda", requires_grad=True)
    x = unrolled_gap_tv(torch.tensor(mask, device="cuda"), tt, lt, wt, logits=False)
    ((x - tt) ** 2).mean().backward()
    glam, gw = lt.grad.cpu().numpy(), wt.grad.cpu().numpy()
    del x
    dd = dict(device="cuda", dtype=torch.float64)
    l64 = torch.tensor(lam, **dd, requires_grad=True)
    w64 = torch.tensor(w, **dd, requires_grad=True)
    R.mse_loss(torch.tensor(mask, **dd), tt.double(), l64, w64, checkpoint=True).backward()
    print("unrolled_batch2_1024x1024x20 lambda/weight gradient errors", rel(glam, l64.grad.cpu().numpy()),
          rel(gw, w64.grad.cpu().numpy()))
    assert rel(glam, l64.grad.cpu().numpy()) <= 1e-3
    assert rel(gw, w64.grad.cpu().numpy()) <= 1e-3


def test_module_and_errors(gpu):
    import torch
    from pnp_sci import _lib
    from pnp_sci.unrolled import UnrolledGAPTV, unrolled_gap_tv
    from pnp_sci.synthetic import unrolled_batch
    truth, z = unrolled_batch(2, 32, 48, 3, seed=5)
    model = UnrolledGAPTV(48, 32, 3, iters=3, logits_init=torch.from_numpy(z).transpose(1, 2))
    x = model(torch.tensor(truth, device="cuda").transpose(2, 3).contiguous())
    ((x - 0.5) ** 2).mean().backward()
    assert model.mask_logits.grad is not None and torch.isfinite(model.mask_logits.grad).all()
    assert model.lambdas.grad.shape == (3,) and model.tv_weights.grad.shape == (3,)
    m = torch.tensor(z, device="cuda")
    t = torch.tensor(truth, device="cuda")
    with pytest.raises(_lib.ScigapError):
        unrolled_gap_tv(m, t, torch.ones(2), torch.full((2,), 0.1), tv_iter_max=9)
    with pytest.raises(_lib.ScigapError):
        unrolled_gap_tv(m, t, torch.ones(2), torch.tensor([0.1, -0.1]))
    with pytest.raises(ValueError):
        unrolled_gap_tv(m, t[:, :2], torch.ones(2), torch.full((2,), 0.1))


def test_batch_lanes_match_single_samples(gpu):
    """Samples of a batch run on up to four concurrent streams with their own
    scratch (batch 6 reuses lanes); each sample's output is bitwise the one a
    batch-1 call gives, and the batch gradients equal the sums of the
    single-sample gradients up to the order of the sums."""
    from pnp_sci.synthetic import unrolled_batch
    B, n = 6, 5
    truth, z = unrolled_batch(B, 72, 136, 4, seed=17)
    lam, w = _schedules(3, 5)
    got = _hip(z, truth, lam, w, n, logits=True)
    xs, gm, gl, gw, loss = [], 0.0, 0.0, 0.0, 0.0
    for b in range(B):
        one = _hip(z, truth[b:b + 1], lam, w, n, logits=True)
        xs.append(one["x"][0])
        # the batch loss is the mean over B samples: scale the per-sample grads
        gm, gl, gw = gm + one["mask"] / B, gl + one["lam"] / B, gw + one["w"] / B
        loss += one["loss"] / B
    assert np.array_equal(got["x"], np.stack(xs))
    assert abs(got["loss"] - loss) <= 1e-6 * loss
    assert rel(got["mask"], gm) <= 1e-5
    assert rel(got["lam"], gl) <= 1e-5 and rel(got["w"], gw) <= 1e-5
