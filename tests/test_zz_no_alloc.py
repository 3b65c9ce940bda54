"""scigap_tv_chambolle_noalloc (scratch inside the library, no workspace
argument) against scigap_tv_chambolle (caller workspace) through the C ABI:
equal results and stop iterations on two calls in a row, the second on a
smaller input, and the documented refusals."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _with_workspace(L, _lib, img, w, eps, n):
    import torch
    T, H, W = img.shape
    out = torch.empty_like(img)
    stop = torch.empty(T, dtype=torch.int32, device="cuda")
    nb = L.scigap_tv_workspace_bytes(H, W, T, n)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    _lib.check(L.scigap_tv_chambolle(img.data_ptr(), out.data_ptr(), H, W, T, w, eps, n,
                                     stop.data_ptr(), ws.data_ptr(), nb, _lib.stream_ptr()))
    return out, stop


def _no_alloc(L, _lib, img, w, eps, n, with_stop=True):
    import torch
    T, H, W = img.shape
    out = torch.empty_like(img)
    stop = torch.empty(T, dtype=torch.int32, device="cuda") if with_stop else None
    _lib.check(L.scigap_tv_chambolle_noalloc(img.data_ptr(), out.data_ptr(), H, W, T, w, eps, n,
                                             stop.data_ptr() if with_stop else None,
                                             _lib.stream_ptr()))
    return out, stop


def test_no_alloc_tv_matches_workspace_tv(gpu):
    import torch
    from pnp_sci import _lib
    L = gpu
    rng = np.random.default_rng(21)
    # a large call, then a smaller one (the scratch keeps the first call's contents)
    cases = [((20, 700, 900), 0.1, 2e-2, 5), ((3, 97, 131), 0.15, 2e-4, 7),
             ((5, 64, 64), 0.2, 5e-2, 3)]
    sizes = []
    for (T, H, W), w, eps, n in cases:
        r, c = np.arange(H)[:, None], np.arange(W)[None, :]
        img = np.stack([0.5 + 0.2 * np.cos(2 * np.pi * (r / H + c / W + t / 7.0)) if t % 2 else
                        rng.random((H, W)) for t in range(T)]).astype(np.float32)
        d = torch.from_numpy(img).cuda()
        want, wstop = _with_workspace(L, _lib, d, w, eps, n)
        got, gstop = _no_alloc(L, _lib, d, w, eps, n)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        assert torch.equal(gstop, wstop)
        got2, _ = _no_alloc(L, _lib, d, w, eps, n, with_stop=False)
        assert torch.equal(got2, want)
        sizes.append(L.scigap_tv_workspace_bytes(H, W, T, n))
    assert sizes[1] < sizes[0]
    # refusals: deep TV, and a shape whose scratch exceeds the library's array
    d = torch.zeros((2, 16, 16), device="cuda")
    o = torch.empty_like(d)
    assert L.scigap_tv_chambolle_noalloc(d.data_ptr(), o.data_ptr(), 16, 16, 2, 0.1, 2e-4, 9,
                                         None, _lib.stream_ptr()) == -6  # SCIGAP_EUNSUPPORTED
    big_T = 2000
    H, W = 2048, 2048
    assert L.scigap_tv_workspace_bytes(H, W, big_T, 5) > L.scigap_tv_noalloc_max_bytes()
    # rejected on the host before any launch: the (small) buffers are never touched
    assert L.scigap_tv_chambolle_noalloc(d.data_ptr(), o.data_ptr(), H, W, big_T, 0.1, 2e-4, 5,
                                         None, _lib.stream_ptr()) == -1  # SCIGAP_EINVAL
    assert b"scratch" in L.scigap_last_error()
