"""CACTI data files as the reference drivers read and write them.

* ``load_cacti_mat`` follows PnP_SCI/python/pnp_sci_demo_kobe.py:52-63 (and
  pnp_sci_test_orig.py:95-110): MATLAB files up to v7.2 are read with
  ``scipy.io.loadmat`` in NumPy order; v7.3 files are HDF5, read with h5py in
  MATLAB (column-major) order and transposed.  Arrays come back as float32.
* ``save_results_mat`` writes the result schema of ``show_n_save_res``
  (PnP_SCI/python/utils.py:98-125): vdenoise, tdenoise, iframe, nframe, Cr,
  plus orig and the PSNR/SSIM series when a ground truth exists.

These are host-side file formats around the solvers; they involve no device
work.
"""

from __future__ import annotations

from statistics import mean

import numpy as np

__all__ = ["load_cacti_mat", "save_results_mat", "mat_version"]


def mat_version(path) -> int:
    """Major MAT-file version: 0 (v4), 1 (v5 .. v7.2) or 2 (v7.3, HDF5)."""
    from scipy.io.matlab import matfile_version
    return int(matfile_version(path)[0])


def load_cacti_mat(path, keys=("meas", "mask", "orig")):
    """{key: float32 array} for the variables present among ``keys``."""
    out = {}
    if mat_version(path) < 2:
        import scipy.io as sio
        f = sio.loadmat(path)
        for k in keys:
            if k in f:
                out[k] = np.float32(f[k])
        return out
    try:
        import h5py
    except ImportError as e:
        raise ImportError("%s is a MATLAB v7.3 (HDF5) file; reading it needs h5py, which is "
                          "not installed" % path) from e
    with h5py.File(path, "r") as f:
        for k in keys:
            if k in f:
                out[k] = np.float32(np.asarray(f[k], order="F")).transpose()
    return out


def save_results_mat(path, vdenoise, tdenoise, psnr_denoise=None, ssim_denoise=None,
                     psnrall_denoise=None, orig=None, iframe=0, nframe=1, Cr=None, **kwargs):
    """Write a results file with the schema of utils.py:98-125 (MATLAB v5)."""
    import scipy.io as sio
    d = {"vdenoise": np.asarray(vdenoise), "tdenoise": tdenoise, "iframe": iframe,
         "nframe": nframe, "Cr": Cr if Cr is not None else np.shape(vdenoise)[-1]}
    if orig is not None:
        d.update({"orig": np.asarray(orig), "psnr_denoise": psnr_denoise,
                  "ssim_denoise": ssim_denoise, "psnrall_denoise": psnrall_denoise,
                  "psnr_mean": mean(psnr_denoise)})
    d.update(kwargs)
    sio.savemat(path, d)
