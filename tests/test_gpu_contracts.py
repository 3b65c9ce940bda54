"""GPU checks of the smaller contracts of the drop-in path.

* Non-finite values in device arrays propagate as in the reference's NumPy
  arithmetic: a NaN in the measurement gives NaN at exactly the pixels where
  the FP64 oracle has NaN, and the other pixels still match it.
* A CACTI file written by scipy (MATLAB v5, the reference's v7.2 branch,
  pnp_sci_demo_kobe.py:52-56) goes through load_cacti_mat and
  admmdenoise_cacti with a warm start v0 and flipped coded frames.
* An unknown projection method prints the reference's message and then fails
  as the reference does (UnboundLocalError, pnp_sci_algo.py:518-529).
"""

import numpy as np
import pytest

import pnp_sci_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("method", ["gap", "admm"])
def test_nan_in_measurement_propagates_like_the_reference(gpu, method):
    import pnp_sci
    from pnp_sci.synthetic import make_problem
    pr = make_problem(48, 64, 4, seed=3)
    y = pr.y.copy()
    y[20, 30] = np.nan
    A, At = pnp_sci.make_operators(pr.mask)
    m = pr.mask.astype(np.float64)
    args = dict(iter_max=2, tv_weight=0.1, show_iqa=False)
    if method == "gap":
        got = pnp_sci.gap_denoise(y, pr.phi_sum, A, At, **args)[0]
        want = O.gap_denoise(y.astype(np.float64), pr.phi_sum.astype(np.float64),
                             lambda v: O.A_(v, m), lambda v: O.At_(v, m), **args)[0]
    else:
        got = pnp_sci.admm_denoise(y, pr.phi_sum, A, At, **args)[0]
        want = O.admm_denoise(y.astype(np.float64), pr.phi_sum.astype(np.float64),
                              lambda v: O.A_(v, m), lambda v: O.At_(v, m), **args)[0]
    nan_want = np.isnan(want)
    assert 0 < nan_want.sum() < nan_want.size
    assert np.array_equal(np.isnan(got), nan_want)
    ok = ~nan_want
    assert rel(got[ok], want[ok]) <= 1e-4


def test_cacti_mat_file_with_warm_start(gpu, tmp_path):
    import scipy.io as sio

    import pnp_sci
    from pnp_sci.io import load_cacti_mat, mat_version
    from pnp_sci.synthetic import truth_video
    rng = np.random.default_rng(9)
    H, W, T, F = 32, 40, 4, 3
    mask = (rng.random((H, W, T)) < 0.5).astype(np.float64)
    orig = truth_video(rng, H, W, T * F) * 255
    meas = np.stack([(orig[..., f * T:(f + 1) * T] * mask).sum(2) for f in range(F)], -1)
    path = str(tmp_path / "scene_cacti.mat")
    sio.savemat(path, {"meas": meas, "mask": mask, "orig": orig})
    assert mat_version(path) == 1
    d = load_cacti_mat(path)
    assert d["meas"].dtype == np.float32 and d["mask"].shape == (H, W, T)
    v0 = (rng.random((H, W, T * F)) * 0.5 + 0.25).astype(np.float32)
    args = dict(denoiser="tv", iter_max=5, tv_weight=0.1, tv_iter_max=5)
    A = lambda v: pnp_sci.A_(v, d["mask"])  # noqa: E731
    At = lambda v: pnp_sci.At_(v, d["mask"])  # noqa: E731
    got = pnp_sci.admmdenoise_cacti(d["meas"], d["mask"], A, At, projmeth="gap", v0=v0,
                                    orig=d["orig"], nframe=F, MAXB=255., maskdirection="updown",
                                    **args)
    m64 = d["mask"].astype(np.float64)
    want = O.admmdenoise_cacti(d["meas"].astype(np.float64), m64, lambda v: O.A_(v, m64),
                               lambda v: O.At_(v, m64), projmeth="gap",
                               v0=v0.astype(np.float64), orig=d["orig"].astype(np.float64),
                               nframe=F, MAXB=255., maskdirection="updown", **args)
    assert got[0].shape == (H, W, T * F) and got[0].dtype == np.float32
    assert rel(got[0], want[0]) <= 1e-4
    assert np.allclose(got[2], want[2], atol=1e-2)    # per-frame PSNR, flipped blocks included
    assert np.allclose(got[3], want[3], atol=1e-4)    # per-frame SSIM
    # block 1 (odd, 'updown') was solved from the reversed warm start
    assert rel(got[0][..., T:2 * T], want[0][..., T:2 * T]) <= 1e-4


def test_unknown_projection_method(gpu, capsys):
    import pnp_sci
    rng = np.random.default_rng(1)
    mask = (rng.random((16, 16, 2)) < 0.5).astype(np.float32)
    meas = (rng.random((16, 16, 1)) * 2).astype(np.float32)
    with pytest.raises(UnboundLocalError):
        pnp_sci.admmdenoise_cacti(meas, mask, lambda v: pnp_sci.A_(v, mask),
                                  lambda v: pnp_sci.At_(v, mask), projmeth="fista",
                                  denoiser="tv", iter_max=2)
    assert "Unsupported projection method FISTA" in capsys.readouterr().out
    with pytest.raises(UnboundLocalError):
        O.admmdenoise_cacti(meas, mask, lambda v: O.A_(v, mask), lambda v: O.At_(v, mask),
                            projmeth="fista", denoiser="tv", iter_max=2)
