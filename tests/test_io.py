"""CACTI .mat files (pnp_sci_demo_kobe.py:52-63, utils.py:98-125): host-side
formats, no device needed."""

import numpy as np
import pytest
import scipy.io as sio

from pnp_sci.io import load_cacti_mat, mat_version, save_results_mat


def test_load_v5_mat_roundtrip(tmp_path):
    g = np.random.default_rng(0)
    meas = g.random((16, 20, 2)) * 255
    mask = (g.random((16, 20, 8)) < 0.5).astype(np.uint8)
    orig = g.random((16, 20, 16)) * 255
    p = tmp_path / "kobe.mat"
    sio.savemat(p, {"meas": meas, "mask": mask, "orig": orig})
    assert mat_version(p) == 1
    d = load_cacti_mat(p)
    assert set(d) == {"meas", "mask", "orig"}
    for k, v in (("meas", meas), ("mask", mask), ("orig", orig)):
        assert d[k].dtype == np.float32 and d[k].shape == v.shape
        assert np.array_equal(d[k], np.float32(v))


def test_v73_file_needs_h5py(tmp_path):
    pytest.importorskip("scipy")
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is installed; the HDF5 path is exercised elsewhere")
    except ImportError:
        pass
    # a MATLAB 7.3 header: 116 bytes of text, 8 of subsystem offset, version 0x0200, 'IM'
    hdr = b"MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: Mon Jan  1 00:00:00 2024 HDF5 schema 1.00 ."
    hdr = hdr.ljust(116, b" ") + b"\x00" * 8 + b"\x00\x02" + b"IM"
    p = tmp_path / "v73.mat"
    p.write_bytes(hdr + b"\x00" * 384)
    assert mat_version(p) == 2
    with pytest.raises(ImportError, match="v7.3"):
        load_cacti_mat(p)


def test_results_schema(tmp_path):
    g = np.random.default_rng(1)
    v = g.random((8, 10, 4)).astype(np.float32)
    o = g.random((8, 10, 4))
    p = tmp_path / "res.mat"
    save_results_mat(p, v, 1.5, [30.0, 31.0, 32.0, 33.0], [0.9, 0.9, 0.9, 0.9],
                     [[20.0, 25.0]], orig=o, iframe=0, nframe=1, Cr=4, extra=7)
    d = sio.loadmat(p)
    for k in ("vdenoise", "orig", "psnr_denoise", "ssim_denoise", "psnrall_denoise",
              "psnr_mean", "tdenoise", "iframe", "nframe", "Cr", "extra"):
        assert k in d, k
    assert np.array_equal(d["vdenoise"], v)
    assert float(d["psnr_mean"].ravel()[0]) == pytest.approx(31.5)
    q = tmp_path / "res_noorig.mat"
    save_results_mat(q, v, 1.5)
    d = sio.loadmat(q)
    assert "orig" not in d and int(d["Cr"].ravel()[0]) == 4
