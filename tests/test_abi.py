"""C ABI checks that need no GPU: the library loads, exports exactly what
include/scigap.h declares, and rejects bad arguments with status codes
before any device work."""

import ctypes
import os
import re

import pytest

from pnp_sci import _lib
from pnp_sci.build import LIB_PATH

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "scigap.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scigap_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def dll():
    if not os.path.exists(LIB_PATH):
        pytest.fail("libscigap.so is not built (run __graft_entry__.build())")
    return _lib.load_library()


def test_header_and_binding_agree():
    decl = declared_functions()
    assert len(decl) >= 14
    assert sorted(_lib.SYMBOLS) == decl


def test_library_exports_every_declared_symbol(dll):
    for name in declared_functions():
        assert hasattr(dll, name), name
    assert dll.scigap_version() >= 10000


def test_shape_and_pointer_validation_without_device(dll):
    null = None
    assert dll.scigap_forward(null, null, null, 0, 4, 4, null) == -1       # EINVAL
    assert b"positive" in dll.scigap_last_error()
    assert dll.scigap_forward(null, null, null, 4, 4, 4, null) == -2       # ENULL
    assert dll.scigap_adjoint(null, null, null, 4, -1, 4, null) == -1
    assert dll.scigap_phi_sum(null, null, 1 << 20, 1 << 20, 1, null) == -1  # too large
    assert dll.scigap_ssim(ctypes.c_void_p(256), ctypes.c_void_p(512), ctypes.c_void_p(1024),
                           5, 9, 1, 1.0, null) == -1                        # H < 7
    assert dll.scigap_tv_chambolle(null, null, 4, 4, 1, 0.1, 2e-4, 5, null, null, 0, null) == -2
    assert dll.scigap_tv_chambolle(ctypes.c_void_p(256), ctypes.c_void_p(4096), 4, 4, 1, 0.1,
                                   2e-4, 0, null, null, 0, null) == -1      # n_iter_max < 1
    # overlapping input/output volumes
    assert dll.scigap_tv_chambolle(ctypes.c_void_p(256), ctypes.c_void_p(260), 4, 4, 1, 0.1,
                                   2e-4, 5, null, ctypes.c_void_p(1 << 20), 1 << 20, null) == -5


def test_workspace_sizes_and_params(dll):
    p = _lib.Params()
    dll.scigap_default_params(p, 256, 256, 8)
    assert (p.H, p.W, p.T, p.iters, p.tv_iter_max) == (256, 256, 8, 50, 5)
    assert p.accelerate == 1 and p.method == _lib.METHOD_GAP
    assert abs(p.gamma - 0.01) < 1e-8 and abs(p.tv_weight - 0.1) < 1e-8  # fp32 fields
    n = dll.scigap_solve_workspace_bytes(p)
    N = 256 * 256
    assert n >= 8 * N * 4 + 4 * N * 4          # x buffer + r, y1 ping-pong
    p.method = _lib.METHOD_ADMM
    assert dll.scigap_solve_workspace_bytes(p) >= n + 8 * N * 4 - 2 * N * 4  # + b buffer
    p.tv_iter_max = 0
    assert dll.scigap_solve_workspace_bytes(p) == 0
    assert b"tv_iter_max" in dll.scigap_last_error()
    assert dll.scigap_tv_workspace_bytes(64, 64, 3, 1) > 0
    assert dll.scigap_tv_workspace_bytes(64, 64, 3, 30) >= 4 * 3 * 64 * 64 * 4
    # solve rejects a missing workspace before touching the device
    p.tv_iter_max = 5
    rc = dll.scigap_solve(p, ctypes.c_void_p(256), ctypes.c_void_p(1 << 22),
                          ctypes.c_void_p(1 << 23), None, ctypes.c_void_p(1 << 24), None, None,
                          None, None, 0, None)
    assert rc == -4 and b"workspace" in dll.scigap_last_error()
    # non-finite scalars are rejected on the host
    for field in ("lambda_", "gamma", "tv_weight", "tv_eps", "tv_weight_decay"):
        q = _lib.Params.from_buffer_copy(p)
        setattr(q, field, float("nan") if field != "tv_weight" else float("inf"))
        assert dll.scigap_solve_workspace_bytes(q) == 0, field
        assert b"finite" in dll.scigap_last_error(), field
    for w, e in ((float("inf"), 2e-4), (0.1, float("nan"))):
        assert dll.scigap_tv_chambolle(ctypes.c_void_p(256), ctypes.c_void_p(4096), 4, 4, 1, w, e,
                                       5, None, None, 0, None) == -1
        assert b"finite" in dll.scigap_last_error()


def test_unrolled_sizes_and_validation(dll):
    u = _lib.Unrolled(H=64, W=48, T=4, batch=2, iters=3, tv_iter_max=5, logits=1, measure=1)
    N, V = 64 * 48, 64 * 48 * 4
    ck = dll.scigap_unrolled_checkpoint_bytes(ctypes.byref(u))
    assert ck >= 2 * (3 * V * 4 + 3 * N * 4 + N * 4)      # x_k, r_k, y per sample
    ws = dll.scigap_unrolled_workspace_bytes(ctypes.byref(u))
    assert ws >= 2 * 2 * V * 4                            # two lanes: vbar x2 each
    for field, bad in (("tv_iter_max", 8), ("tv_iter_max", 1), ("iters", 0), ("batch", 0)):
        v = _lib.Unrolled.from_buffer_copy(u)
        setattr(v, field, bad)
        assert dll.scigap_unrolled_checkpoint_bytes(ctypes.byref(v)) == 0
        assert dll.scigap_unrolled_workspace_bytes(ctypes.byref(v)) == 0
    odd = _lib.Unrolled(H=5, W=7, T=3, batch=2, iters=2, tv_iter_max=3, logits=0, measure=0)
    assert dll.scigap_unrolled_checkpoint_bytes(ctypes.byref(odd)) == 0
    odd.batch = 1
    assert dll.scigap_unrolled_checkpoint_bytes(ctypes.byref(odd)) > 0
    # host-side checks run before any device work
    lam = (ctypes.c_float * 3)(1, 1, 1)
    w = (ctypes.c_float * 3)(0.1, 0.0, 0.1)
    big = ctypes.c_void_p(1 << 30)
    rc = dll.scigap_unrolled_forward(ctypes.byref(u), lam, w, ctypes.c_void_p(256),
                                     ctypes.c_void_p(1 << 20), ctypes.c_void_p(1 << 24), big, ck,
                                     ctypes.c_void_p(1 << 32), ws, None)
    assert rc == -1 and b"weight" in dll.scigap_last_error()
    w[1] = 0.1
    rc = dll.scigap_unrolled_forward(ctypes.byref(u), lam, w, ctypes.c_void_p(256),
                                     ctypes.c_void_p(1 << 20), ctypes.c_void_p(1 << 24), big,
                                     ck - 1, ctypes.c_void_p(1 << 32), ws, None)
    assert rc == -4 and b"checkpoint" in dll.scigap_last_error()
    rc = dll.scigap_unrolled_backward(ctypes.byref(u), lam, w, ctypes.c_void_p(256),
                                      ctypes.c_void_p(1 << 20), big, ck, ctypes.c_void_p(1 << 24),
                                      None, ctypes.c_void_p(1 << 25), None, None,
                                      ctypes.c_void_p(1 << 32), ws, None)
    assert rc == -1 and b"truth" in dll.scigap_last_error()   # measure=1 gives no dL/dtruth
