"""Pin the FP64 C restatement against the NumPy oracle (CPU only)."""

import numpy as np
import pytest

import pnp_sci_oracle as O

oracle_c = pytest.importorskip("oracle_c", reason="make -C oracle builds the C oracle")


def test_c_tv_matches_numpy_oracle():
    rng = np.random.default_rng(0)
    img = rng.random((33, 41, 3))
    for n, eps in ((1, 2e-4), (2, 2e-4), (5, 2e-4), (30, 2e-4), (30, 5e-2)):
        out, ret = oracle_c.tv(img, 0.15, eps, n)
        for t in range(3):
            st = {}
            want = O.tv_chambolle_nd(img[..., t], 0.15, eps, n, _stats=st)
            assert np.max(np.abs(out[..., t] - want)) < 1e-12
            assert ret[t] == st["returned"]


@pytest.mark.parametrize("method,kw", [("gap", {}), ("gap_plain", {}), ("admm", {}),
                                       ("admm", {"clip": True, "gamma": 0.0}),
                                       ("admm", {"weight_decay": 0.999, "gamma_decay": 0.998})])
def test_c_solver_matches_numpy_oracle(method, kw):
    rng = np.random.default_rng(1)
    H, W, T = 24, 30, 5
    mask = (rng.random((H, W, T)) < 0.5).astype(np.float64)
    truth = rng.random((H, W, T))
    y = (truth * mask).sum(2)
    ps = O.phi_sum(mask)
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    got, sq = oracle_c.gap_tv(y, mask, ps, iters=7, method=method, tv_weight=0.1, truth=truth,
                              **kw)
    if method.startswith("gap"):
        want = O.gap_denoise(y, ps, A, At, accelerate=(method == "gap"), iter_max=7,
                             tv_weight=0.1, show_iqa=False)[0]
    elif "weight_decay" in kw:
        want = O.ADMM_TV_rec(y, mask, lambda f, m: O.A_(f, m), lambda v, m: O.At_(v, m), ps, 7,
                             1.0, 0.1, H, W, T, 0.01, None)
        got = oracle_c.gap_tv(y, mask, ps, iters=7, method="admm", tv_weight=0.1,
                              tv_iter_max=30, **kw)
    else:
        want = O.admm_denoise(y, ps, A, At, iter_max=7, tv_weight=0.1, show_iqa=False,
                              clip=kw.get("clip", False), gamma=kw.get("gamma", 0.01))[0]
    assert np.max(np.abs(got - want)) < 1e-10


@pytest.mark.parametrize("method", ["gap", "admm"])
def test_c_solver_exit_counts_match_numpy_oracle(method, monkeypatch):
    """The per-iteration early-stop counts of the C oracle equal those of the
    NumPy oracle's TV calls (eps loose enough that some frames stop)."""
    rng = np.random.default_rng(5)
    H, W, T, K = 20, 26, 6, 6
    mask = (rng.random((H, W, T)) < 0.5).astype(np.float64)
    truth = rng.random((H, W, T))
    y = (truth * mask).sum(2)
    ps = O.phi_sum(mask)
    calls = []
    tv64 = O.tv_chambolle_nd

    def counting(image, weight=0.1, eps=2e-4, n_iter_max=200, _stats=None):
        st = {}
        res = tv64(image, weight, eps, n_iter_max, _stats=st)
        calls.append(st["returned"] < n_iter_max - 1)
        return res
    monkeypatch.setattr(O, "tv_chambolle_nd", counting)
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    run = O.gap_denoise if method == "gap" else O.admm_denoise
    run(y, ps, A, At, iter_max=K, tv_weight=0.1, show_iqa=False, _tv_eps=2e-2)
    want = np.array(calls).reshape(K, T).sum(1)
    _, ex = oracle_c.gap_tv(y, mask, ps, iters=K, method=method, tv_weight=0.1, eps=2e-2,
                            exits=True)
    assert 0 < want.sum() < K * T
    assert ex.tolist() == want.tolist()
