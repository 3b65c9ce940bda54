"""Pins of the plain-PyTorch unrolled GAP-TV reference (oracle/unrolled_ref.py)
against the NumPy oracle, on the CPU in float64.

The forward pass must equal the oracle's gap_denoise (itself pinned to
scikit-image goldens) with the energy exit off, and autograd's gradients must
equal central finite differences of the NumPy oracle's forward pass."""

import numpy as np
import pytest
import torch

import pnp_sci_oracle as O
import unrolled_ref as R

H, W, T, B, K = 12, 10, 3, 2, 3


def _problem(seed=0):
    g = np.random.default_rng(seed)
    truth = g.random((B, T, H, W))
    z = g.standard_normal((T, H, W))
    lam = 1.0 + 0.3 * g.standard_normal(K)
    w = 0.05 + 0.1 * g.random(K)
    return truth, z, lam, w


def _sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def _oracle_x(phi_thw, y_hw, lam, w, n=5):
    """Accelerated GAP-TV from the oracle's primitives in the reference's loop
    order (pnp_sci_algo.py:638-650), one lambda/weight per iteration,
    TV energy exit off.  phi (T,H,W), y (H,W); returns (T,H,W)."""
    m = np.moveaxis(phi_thw, 0, -1)
    S = O.phi_sum(m)
    x = O.At_(y_hw, m)
    y1 = np.zeros_like(y_hw)
    for k in range(len(lam)):
        yb = O.A_(x, m)
        y1 = y1 + (y_hw - yb)
        x = x + lam[k] * O.At_((y1 - yb) / S, m)
        x = O.denoise_tv_chambolle(x, w[k], eps=0.0, n_iter_max=n, multichannel=True)
    return np.moveaxis(x, -1, 0)


def _oracle_loss(truth, z, lam, w, n=5):
    phi = _sigmoid(z)
    tot = 0.0
    for b in range(B):
        y = (phi * truth[b]).sum(0)
        x = _oracle_x(phi, y, lam, w, n)
        tot += ((x - truth[b]) ** 2).sum()
    return tot / truth.size


def test_tv_fixed_matches_oracle_tv():
    g = np.random.default_rng(3)
    v = g.random((2, 9, 13))
    for n in (1, 2, 5, 8):
        got = R.tv_fixed(torch.from_numpy(v), 0.07, n).numpy()
        for f in range(2):
            want = O.tv_chambolle_nd(v[f], 0.07, eps=0.0, n_iter_max=n)
            assert np.max(np.abs(got[f] - want)) <= 1e-13


def test_forward_constant_schedule_matches_gap_denoise():
    truth, z, _, _ = _problem()
    phi = (np.random.default_rng(1).random((T, H, W)) > 0.5).astype(np.float64)
    for n in (2, 5):
        x = R.unrolled_gap_tv(torch.from_numpy(phi), R.measure(torch.from_numpy(phi),
                                                               torch.from_numpy(truth)),
                              [1.0] * K, [0.1] * K, tv_iter_max=n).numpy()
        for b in range(B):
            m = np.moveaxis(phi, 0, -1)
            y = O.A_(np.moveaxis(truth[b], 0, -1), m)
            want = O.gap_denoise(y, O.phi_sum(m), lambda v: O.A_(v, m), lambda v: O.At_(v, m),
                                 iter_max=K, tv_weight=0.1, tv_iter_max=n, show_iqa=False,
                                 _tv_eps=0.0)[0]
            assert np.max(np.abs(np.moveaxis(x[b], 0, -1) - want)) <= 1e-12


def test_forward_schedule_and_sigmoid_mask():
    truth, z, lam, w = _problem(2)
    phi = _sigmoid(z)
    x = R.unrolled_gap_tv(torch.from_numpy(phi), R.measure(torch.from_numpy(phi),
                                                           torch.from_numpy(truth)),
                          torch.from_numpy(lam), torch.from_numpy(w)).numpy()
    for b in range(B):
        want = _oracle_x(phi, (phi * truth[b]).sum(0), lam, w)
        assert np.max(np.abs(x[b] - want)) <= 1e-12


def test_autograd_matches_oracle_finite_differences():
    truth, z, lam, w = _problem(4)
    zt = torch.from_numpy(z).requires_grad_()
    lt = torch.from_numpy(lam).requires_grad_()
    wt = torch.from_numpy(w).requires_grad_()
    loss = R.mse_loss(torch.sigmoid(zt), torch.from_numpy(truth), lt, wt)
    assert abs(loss.item() - _oracle_loss(truth, z, lam, w)) <= 1e-14 * max(loss.item(), 1)
    loss.backward()
    g = np.concatenate([zt.grad.numpy().ravel(), lt.grad.numpy(), wt.grad.numpy()])
    rng = np.random.default_rng(5)
    eps = 1e-6
    nz = z.size
    for _ in range(6):
        d = rng.standard_normal(g.size)
        d /= np.linalg.norm(d)
        dz, dl, dw = d[:nz].reshape(z.shape), d[nz:nz + K], d[nz + K:]
        fp = _oracle_loss(truth, z + eps * dz, lam + eps * dl, w + eps * dw)
        fm = _oracle_loss(truth, z - eps * dz, lam - eps * dl, w - eps * dw)
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd) + 1e-12, (fd, g @ d)


def test_autograd_matches_finite_differences_in_y():
    truth, z, lam, w = _problem(6)
    phi = _sigmoid(z)
    y0 = np.stack([(phi * truth[b]).sum(0) for b in range(B)])
    yt = torch.from_numpy(y0.copy()).requires_grad_()
    x = R.unrolled_gap_tv(torch.from_numpy(phi), yt, torch.from_numpy(lam), torch.from_numpy(w))
    loss = ((x - torch.from_numpy(truth)) ** 2).mean()
    loss.backward()

    def f(y):
        return sum(((_oracle_x(phi, y[b], lam, w) - truth[b]) ** 2).sum()
                   for b in range(B)) / truth.size

    d = np.random.default_rng(7).standard_normal(y0.shape)
    d /= np.linalg.norm(d)
    eps = 1e-6
    fd = (f(y0 + eps * d) - f(y0 - eps * d)) / (2 * eps)
    assert abs(fd - float((yt.grad.numpy() * d).sum())) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("n", [2, 3])
def test_checkpointed_reference_equals_plain(n):
    truth, z, lam, w = _problem(8)
    phi = torch.sigmoid(torch.from_numpy(z))
    grads = []
    for ck in (False, True):
        lt = torch.from_numpy(lam).requires_grad_()
        wt = torch.from_numpy(w).requires_grad_()
        R.mse_loss(phi, torch.from_numpy(truth), lt, wt, tv_iter_max=n,
                   checkpoint=ck).backward()
        grads.append(torch.cat([lt.grad, wt.grad]).numpy())
    assert np.max(np.abs(grads[0] - grads[1])) <= 1e-14
