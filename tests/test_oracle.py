"""Pins for the NumPy oracle (CPU only).

The oracle is checked against (1) golden vectors written by scikit-image
0.18.3 itself (tests/golden/make_tv_golden.py), bit-exact in float32 and
float64, and (2) properties the mathematics fixes, chosen so that a dropped
term, a wrong sign or a shifted index fails at least one of them.
"""

import numpy as np
import pytest
import scipy.ndimage as ndi

import pnp_sci_oracle as O


def _golden(golden_dir, name):
    return np.load("%s/%s" % (golden_dir, name), allow_pickle=False)


# ---------------------------------------------------------------------------
# Chambolle TV against scikit-image 0.18.3
# ---------------------------------------------------------------------------

def test_tv_matches_skimage_golden_bit_exact(golden_dir):
    z = _golden(golden_dir, "tv_chambolle_golden.npz")
    assert str(z["skimage_version"]) == "0.18.3"
    meta = z["meta"]
    assert len(meta) > 100
    for k, (w, n, eps) in enumerate(meta):
        img, want = z["in_%d" % k], z["out_%d" % k]
        got = O.denoise_tv_chambolle(img, float(w), eps=float(eps), n_iter_max=int(n),
                                     multichannel=True)
        assert got.dtype == want.dtype
        assert np.array_equal(got, want), "case %d (%s, w=%g, n=%d)" % (k, img.shape, w, n)


def test_gap_admm_loops_match_golden_bit_exact(golden_dir):
    g = _golden(golden_dir, "gaptv_small_golden.npz")
    for k in range(len(g["meta"])):
        y, m = g["y_%d" % k], g["mask_%d" % k]
        A = lambda x: O.A_(x, m)  # noqa: E731
        At = lambda v: O.At_(v, m)  # noqa: E731
        ps = O.phi_sum(m)
        runs = {
            "gap_acc": O.gap_denoise(y, ps, A, At, iter_max=12, tv_weight=0.1, show_iqa=False)[0],
            "gap_plain": O.gap_denoise(y, ps, A, At, accelerate=False, iter_max=12,
                                       tv_weight=0.1, show_iqa=False)[0],
            "admm": O.admm_denoise(y, ps, A, At, iter_max=12, tv_weight=0.1, show_iqa=False)[0],
            "admm_clip": O.admm_denoise_clip(y, ps, A, At, iter_max=12, tv_weight=0.1,
                                             show_iqa=False)[0],
        }
        for name, got in runs.items():
            assert np.array_equal(got, g["%s_%d" % (name, k)]), (name, k)


def test_tv_early_exit_branch_is_exercised():
    # A smooth frame stops long before n_iter_max; the golden cases above pin
    # what is returned, this pins that the branch is reached at all.
    r = np.arange(48)[:, None]
    c = np.arange(40)[None, :]
    img = 0.5 + 0.1 * np.cos(2 * np.pi * (r / 48.0 + c / 80.0))
    it = {}
    out = O.tv_chambolle_nd(img, 0.1, 2e-4, 200, _stats=it)
    assert 1 <= it["returned"] < 199
    # The same frame with a tiny eps runs to the end.
    it2 = {}
    O.tv_chambolle_nd(img, 0.1, 1e-30, 12, _stats=it2)
    assert it2["returned"] == 11
    assert out.shape == img.shape


# ---------------------------------------------------------------------------
# Chambolle TV: properties
# ---------------------------------------------------------------------------

def test_tv_one_iteration_is_identity():
    img = np.random.default_rng(0).random((9, 7))
    assert np.array_equal(O.tv_chambolle_nd(img, 0.3, n_iter_max=1), img)


def test_tv_constant_frame_unchanged():
    img = np.full((6, 5), 0.37)
    assert np.array_equal(O.tv_chambolle_nd(img, 0.2, n_iter_max=7), img)


@pytest.mark.parametrize("n", [2, 3, 5, 20])
def test_tv_preserves_frame_mean(n):
    img = np.random.default_rng(n).random((13, 11))
    out = O.tv_chambolle_nd(img, 0.2, n_iter_max=n)
    assert abs(out.mean() - img.mean()) < 1e-14


def test_tv_large_weight_is_neumann_heat_step():
    # n=2 performs one dual step p = -tau*g/(1+(tau/w)|g|); for w -> inf the
    # result is img + tau*Laplacian(img) with Neumann (replicated) borders,
    # computed independently by scipy.ndimage.laplace.
    img = np.random.default_rng(1).random((12, 17))
    got = O.tv_chambolle_nd(img, 1e12, n_iter_max=2)
    want = img + 0.25 * ndi.laplace(img, mode="nearest")
    assert np.max(np.abs(got - want)) < 1e-10


def test_tv_frames_independent_and_tv_decreases():
    rng = np.random.default_rng(2)
    vol = rng.random((20, 18, 3))
    out = O.denoise_tv_chambolle(vol, 0.1, n_iter_max=5, multichannel=True)
    for c in range(3):
        assert np.array_equal(out[..., c], O.tv_chambolle_nd(vol[..., c], 0.1, n_iter_max=5))

    def tv(a):
        return np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum()

    for c in range(3):
        assert tv(out[..., c]) < tv(vol[..., c])
        assert out[..., c].std() < vol[..., c].std()


# ---------------------------------------------------------------------------
# Forward model and solvers: properties
# ---------------------------------------------------------------------------

def _problem(H=10, W=9, T=4, seed=0, p=0.5):
    rng = np.random.default_rng(seed)
    mask = (rng.random((H, W, T)) < p).astype(np.float64)
    x = rng.random((H, W, T))
    return mask, x, rng


def test_forward_adjoint_pair():
    mask, x, rng = _problem()
    y = rng.random(mask.shape[:2])
    assert np.isclose(np.sum(O.A_(x, mask) * y), np.sum(x * O.At_(y, mask)), rtol=1e-13)
    assert np.array_equal(O.At_(np.ones(mask.shape[:2]), mask), mask)
    assert np.allclose(O.A_(O.At_(y, mask), mask), y * (mask ** 2).sum(axis=2), rtol=1e-14)


def test_phi_sum_zero_fix():
    mask, _, _ = _problem()
    mask[3, 4, :] = 0
    s = O.phi_sum(mask)
    assert s[3, 4] == 1
    assert np.array_equal(s[mask.sum(2) > 0], mask.sum(2)[mask.sum(2) > 0])


def test_plain_gap_identity_denoiser_is_exact_projection(capsys):
    mask, x, _ = _problem(T=5)
    mask[2, 2, :] = 0  # a pixel no frame sees: Phi_sum fixed to 1, y = 0 there
    y = O.A_(x, mask)
    ps = O.phi_sum(mask)
    xr, _, _, _ = O.gap_denoise(y, ps, lambda v: O.A_(v, mask), lambda v: O.At_(v, mask),
                                accelerate=False, iter_max=1, tvm="none", show_iqa=False)
    assert np.allclose(O.A_(xr, mask), y, atol=1e-12)
    assert "no such tv denoiser" in capsys.readouterr().out


def test_accelerated_first_step_from_zero_equals_plain_step():
    mask, x, _ = _problem(T=6, seed=3)
    y = O.A_(x, mask)
    ps = O.phi_sum(mask)
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    x0 = np.zeros_like(x)
    a = O.gap_denoise(y, ps, A, At, accelerate=True, iter_max=1, tv_iter_max=3, x0=x0,
                      show_iqa=False)[0]
    b = O.gap_denoise(y, ps, A, At, accelerate=False, iter_max=1, tv_iter_max=3, x0=x0,
                      show_iqa=False)[0]
    assert np.array_equal(a, b)
    c = O.gap_denoise(y, ps, A, At, accelerate=True, iter_max=2, tv_iter_max=3, x0=x0,
                      show_iqa=False)[0]
    d = O.gap_denoise(y, ps, A, At, accelerate=False, iter_max=2, tv_iter_max=3, x0=x0,
                      show_iqa=False)[0]
    assert not np.array_equal(c, d)  # the y1 memory matters from step 2 on


def test_admm_gamma0_identity_equals_plain_gap():
    mask, x, _ = _problem(T=4, seed=4)
    y = O.A_(x, mask)
    ps = O.phi_sum(mask)
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    a = O.admm_denoise(y, ps, A, At, gamma=0.0, iter_max=6, tv_iter_max=1, show_iqa=False)[0]
    b = O.gap_denoise(y, ps, A, At, accelerate=False, iter_max=6, tv_iter_max=1,
                      show_iqa=False)[0]
    assert np.array_equal(a, b)


def test_gap_psnr_trajectory_and_errors():
    mask, x, _ = _problem(H=16, W=16, T=4, seed=5)
    y = O.A_(x, mask)
    ps = O.phi_sum(mask)
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    xr, p_, _, pall = O.gap_denoise(y, ps, A, At, iter_max=[3, 2], sigma=[None, None],
                                    X_orig=x, show_iqa=True)
    assert len(pall) == 5 and len(p_) == 4
    assert pall[-1] == pytest.approx(O.psnr(x, xr))
    with pytest.raises(ValueError):
        O.gap_denoise(y, ps, A, At, denoiser="bm3d")
    with pytest.raises(NameError):
        O.gap_denoise(y, ps, A, At, tvm="ITV2D_cham", iter_max=1)
    assert O.psnr(x, x) == 100


def test_cacti_driver_flips_and_normalises():
    rng = np.random.default_rng(6)
    H, W, T, F = 12, 10, 4, 2
    mask = (rng.random((H, W, T)) < 0.5).astype(np.float32)
    orig = rng.random((H, W, T * F)).astype(np.float32) * 255
    meas = np.stack([O.A_(orig[..., f * T:(f + 1) * T], mask) for f in range(F)], -1)
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    args = dict(denoiser="tv", iter_max=3, tv_weight=0.1, tv_iter_max=3, show_iqa=False)
    x_, _, p_, _, pall = O.admmdenoise_cacti(meas, mask, A, At, projmeth="gap", orig=orig,
                                             nframe=F, MAXB=255., maskdirection="updown", **args)
    assert x_.shape == (H, W, T * F) and x_.dtype == np.float32
    # Frame block 0 is 'up' (no flip): identical to a direct solve.
    direct = O.gap_denoise(meas[..., 0] / 255., O.phi_sum(mask), A, At, **args)[0]
    assert np.allclose(x_[..., :T], direct, rtol=1e-6)
    assert len(p_) == T * F and len(pall) == F


def test_ssim_matches_skimage_golden(golden_dir):
    z = _golden(golden_dir, "ssim_golden.npz")
    for k, want in enumerate(z["ssim"]):
        got = O.compare_ssim(z["x_%d" % k], z["y_%d" % k], data_range=1.)
        assert abs(got - want) < 1e-12, (k, got, want)
    with pytest.raises(ValueError):
        O.compare_ssim(np.zeros((6, 9)), np.zeros((6, 9)))


def test_bayer_gap_equals_four_subgrid_solves():
    """gap_denoise_bayer (pnp_sci_algo.py:20-265) decouples into four GAP-TV
    solves on the 2x2 sub-grids: the GAP update is per sub-grid and the
    multichannel TV is per frame."""
    g = np.random.default_rng(11)
    H, W, T = 24, 30, 3
    truth = g.random((H, W, T))
    mask = (g.random((H, W, T)) < 0.5).astype(np.float64)
    y = O.A_(truth, mask)
    x, psnr_, ssim_, psnr_all = O.gap_denoise_bayer(y, mask, iter_max=4, tv_weight=0.1,
                                                    X_orig=truth)
    want = np.zeros_like(mask)
    for b0, b1 in [[0, 0], [0, 1], [1, 0], [1, 1]]:
        m = mask[b0::2, b1::2]
        ps = m.sum(2)
        ps[ps == 0] = 1
        xc = O.gap_denoise(y[b0::2, b1::2], ps, lambda v: O.A_(v, m), lambda v: O.At_(v, m),
                           iter_max=4, tv_weight=0.1, show_iqa=False)[0]
        want[b0::2, b1::2] = xc
    assert np.array_equal(x, want)
    mse = np.mean((x - truth) ** 2)
    assert abs(psnr_all[-1] - 10 * np.log10(1 / mse)) < 1e-9
    assert len(psnr_all) == 4 and len(psnr_) == T and len(ssim_) == T


# ---------------------------------------------------------------------------
# Legacy solvers, PSNR and the coded-frame flips
# ---------------------------------------------------------------------------

def test_legacy_solvers_match_skimage_golden_bit_exact(golden_dir):
    """GAP_TV_rec / ADMM_TV_rec (pnp_sci_algo.py:866-907) against fixtures
    computed around scikit-image's own 30-iteration TV, with the 0.999/0.998
    decays; float32 and float64."""
    z = _golden(golden_dir, "legacy_golden.npz")
    for k, (h, w, t, iters, wt, eta) in enumerate(z["meta"]):
        h, w, t, iters = int(h), int(w), int(t), int(iters)
        y, m = z["y_%d" % k], z["mask_%d" % k]
        ps = O.phi_sum(m)
        g = O.GAP_TV_rec(y, m, O.A_, O.At_, ps, iters, 1.0, float(wt), h, w, t, None)
        a = O.ADMM_TV_rec(y, m, O.A_, O.At_, ps, iters, 1.0, float(wt), h, w, t, float(eta), None)
        assert np.array_equal(g, z["gap_%d" % k]), ("gap", k)
        assert np.array_equal(a, z["admm_%d" % k]), ("admm", k)


def test_legacy_fixtures_exercise_exit_and_decay(golden_dir):
    """The fixtures above only pin what they exercise: some 30-iteration TV
    calls stop early on the energy test, and leaving out the decays (or
    applying them before the first iteration) changes the ADMM result."""
    z = _golden(golden_dir, "legacy_golden.npz")
    k = 3  # 17x31x5, 25 iterations, float64
    h, w, t, iters, wt, eta = z["meta"][k]
    h, w, t, iters = int(h), int(w), int(t), int(iters)
    y, m = z["y_%d" % k], z["mask_%d" % k]
    ps = O.phi_sum(m)
    want = z["admm_%d" % k]
    # no decay: a plain ADMM loop with the same TV depth
    nodecay = O.admm_denoise(y, ps, lambda v: O.A_(v, m), lambda v: O.At_(v, m),
                             gamma=float(eta), iter_max=iters, tv_weight=float(wt),
                             tv_iter_max=30, show_iqa=False)[0]
    assert not np.allclose(nodecay, want, rtol=0, atol=1e-6)
    # decays applied before the first use
    early = O.ADMM_TV_rec(y, m, O.A_, O.At_, ps, iters, 1.0, 0.999 * float(wt), h, w, t,
                          0.998 * float(eta), None)
    assert not np.allclose(early, want, rtol=0, atol=1e-6)
    # 30-iteration TV calls on these frames stop early on the energy test for
    # some frames and run all 30 iterations for others
    returned = []
    for f in range(t):
        st = {}
        O.tv_chambolle_nd(want[..., f], float(wt), n_iter_max=30, _stats=st)
        returned.append(st["returned"])
    assert min(returned) < 29 <= max(returned), returned


def test_legacy_gap_equals_generic_solver(golden_dir):
    """GAP_TV_rec is gap_denoise (accelerated, lambda = step_size) with a
    30-iteration TV and x0 = At(y): two code paths of the oracle meet the
    same float64 fixtures.  (In float32 the legacy loop's y1 = zeros((row,
    col)) is float64 and promotes the update, so the two differ there.)"""
    z = _golden(golden_dir, "legacy_golden.npz")
    for k, (h, w, t, iters, wt, eta) in enumerate(z["meta"]):
        y, m = z["y_%d" % k], z["mask_%d" % k]
        if y.dtype != np.float64:
            continue
        got = O.gap_denoise(y, O.phi_sum(m), lambda v: O.A_(v, m), lambda v: O.At_(v, m),
                            iter_max=int(iters), tv_weight=float(wt), tv_iter_max=30,
                            show_iqa=False)[0]
        assert np.array_equal(got, z["gap_%d" % k]), k


def test_psnr_matches_skimage_golden(golden_dir):
    """compare_psnr against scikit-image's peak_signal_noise_ratio values, and
    psnr (utils.py:28-36, 20*log10(1/sqrt(mse)) in the input dtype) against
    the same values: equal in float64, within float32 rounding otherwise."""
    z = _golden(golden_dir, "psnr_golden.npz")
    for k, want in enumerate(z["psnr"]):
        ref, img = z["ref_%d" % k], z["img_%d" % k]
        assert abs(O.compare_psnr(ref, img, data_range=1.) - want) < 1e-12, k
        tol = 1e-10 if ref.dtype == np.float64 else 1e-4
        assert abs(O.psnr(ref, img) - want) < tol, (k, O.psnr(ref, img), want)


def test_psnr_closed_forms():
    # mse = 0.01 exactly -> 20 dB; mse = 1e-4 -> 40 dB; identical -> 100 (utils)
    # and inf (scikit-image); data_range scales as 20*log10(range).
    ref = np.zeros((4, 5))
    img = np.full((4, 5), 0.1)
    assert O.psnr(ref, img) == pytest.approx(20.0, abs=1e-12)
    assert O.psnr(ref, img / 10) == pytest.approx(40.0, abs=1e-12)
    assert O.psnr(ref, ref) == 100
    assert O.compare_psnr(ref, ref) == np.inf
    assert O.compare_psnr(ref, img, data_range=255.) == pytest.approx(
        20.0 + 20 * np.log10(255.), abs=1e-9)
    # the mean runs over every element, not one frame
    img3 = np.zeros((4, 5, 2))
    img3[..., 1] = 0.1
    assert O.psnr(np.zeros((4, 5, 2)), img3) == pytest.approx(10 * np.log10(200.), abs=1e-12)


@pytest.mark.parametrize("direction", ["updown", "downup"])
def test_cacti_flips_odd_even_blocks(direction):
    """admmdenoise_cacti (pnp_sci_algo.py:504-510,521-526): under 'updown' the
    odd coded frames, under 'downup' the even ones, solve with the warm start
    reversed along frames and return the result reversed (with its per-frame
    PSNR/SSIM lists); the mask and the truth are not reversed."""
    rng = np.random.default_rng(8)
    H, W, T, F = 10, 12, 4, 3
    mask = (rng.random((H, W, T)) < 0.5).astype(np.float64)
    orig = rng.random((H, W, T * F)) * 255
    meas = np.stack([O.A_(orig[..., f * T:(f + 1) * T], mask) for f in range(F)], -1)
    v0 = rng.random((H, W, T * F))
    A = lambda v: O.A_(v, mask)  # noqa: E731
    At = lambda v: O.At_(v, mask)  # noqa: E731
    args = dict(denoiser="tv", iter_max=3, tv_weight=0.1, tv_iter_max=3, show_iqa=False)
    x_, _, p_, s_, pall = O.admmdenoise_cacti(meas, mask, A, At, projmeth="admm", v0=v0,
                                              orig=orig, nframe=F, MAXB=255.,
                                              maskdirection=direction, **args)
    ps = O.phi_sum(mask)
    for kf in range(F):
        flip = (kf % 2 == 1) if direction == "updown" else (kf % 2 == 0)
        v0k = v0[..., kf * T:(kf + 1) * T]
        xk, pk, sk, ak = O.admm_denoise(meas[..., kf] / 255., ps, A, At,
                                        x0=v0k[..., ::-1] if flip else v0k,
                                        X_orig=orig[..., kf * T:(kf + 1) * T] / 255., **args)
        if flip:
            xk, pk, sk, ak = xk[..., ::-1], pk[::-1], sk[::-1], ak[::-1]
        assert np.array_equal(x_[..., kf * T:(kf + 1) * T], xk.astype(np.float32)), kf
        assert p_[kf * T:(kf + 1) * T] == list(pk) and s_[kf * T:(kf + 1) * T] == list(sk)
        assert list(pall[kf]) == list(ak)
        # a flipped block differs from the unflipped solve (the flip is not a no-op)
        if flip:
            plain = O.admm_denoise(meas[..., kf] / 255., ps, A, At, x0=v0k, **args)[0]
            assert not np.allclose(plain, xk[..., ::-1])
    # iframe shifts the parity: block 0 of iframe=1 is an odd coded frame
    x1 = O.admmdenoise_cacti(meas, mask, A, At, projmeth="admm", v0=v0[..., T:2 * T],
                             nframe=1, iframe=1, MAXB=255., maskdirection=direction, **args)[0]
    k1 = O.admm_denoise(meas[..., 1] / 255., ps, A, At,
                        x0=(v0[..., T:2 * T][..., ::-1] if direction == "updown"
                            else v0[..., T:2 * T]), **args)[0]
    want = k1[..., ::-1] if direction == "updown" else k1
    assert np.array_equal(x1, want.astype(np.float32))


def test_borderline_case_fp32_drift():
    """The 9x7x1030 solve (seed 4, 4 iterations, default eps) has exit tests
    close to their threshold: the reference's float32 arithmetic (this
    oracle in float32) and FP64 decide one TV call differently and land
    1.2e-3 apart.  The GPU test of the same case is held to a multiple of
    this drift (tests/test_gpu_configs.py)."""
    from pnp_sci.synthetic import make_problem
    H, W, T, K = 9, 7, 1030, 4
    pr = make_problem(H, W, T, seed=4)
    res = {}
    for dt in (np.float32, np.float64):
        m = pr.mask.astype(dt)
        res[dt] = O.gap_denoise(pr.y.astype(dt), pr.phi_sum.astype(dt), lambda v: O.A_(v, m),
                                lambda v: O.At_(v, m), iter_max=K, tv_weight=0.1,
                                show_iqa=False)[0]
    x32, x64 = res[np.float32].astype(np.float64), res[np.float64]
    drift = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    assert 1.0e-3 < drift < 1.5e-3, drift


def test_tv_exit_margin_reaches_zero_where_the_decision_flips():
    """The stats hook's ``margin`` (distance of the closest exit test from its
    threshold) goes to zero exactly where a change of eps changes the returned
    iteration: bisecting eps across such a change leaves both ends with a
    margin that shrinks with the bracket."""
    rng = np.random.default_rng(4)
    img = rng.random((12, 9))

    def returned(eps):
        st = {}
        O.tv_chambolle_nd(img, 0.1, eps, 30, _stats=st)
        return st["returned"], st["margin"]
    lo, hi = 1e-5, 1e-1
    r_lo, r_hi = returned(lo)[0], returned(hi)[0]
    assert r_lo > r_hi  # a larger eps stops earlier
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if returned(mid)[0] == r_lo:
            lo = mid
        else:
            hi = mid
    assert returned(lo)[1] < 1e-9 and returned(hi)[1] < 1e-9
    assert returned(1e-5)[1] > 1e-3
