"""MI355X-native GAP-TV / ADMM-TV reconstruction for video snapshot
compressive imaging: the ``denoiser='tv'`` path of PnP-SCI
(PnP_SCI/python/pnp_sci_algo.py, utils.py) on hand-written gfx950 kernels.

Public names mirror the reference modules:

* ``pnp_sci.pnp_sci_algo``: gap_denoise, admm_denoise, admmdenoise_cacti,
  GAP_TV_rec, ADMM_TV_rec, gap_denoise_bayer
* ``pnp_sci.utils``: A_, At_, psnr, make_operators
* ``pnp_sci.tv``: denoise_tv_chambolle (scikit-image signature)

Beyond the reference:

* ``pnp_sci.batch``: BatchPlan / gap_denoise_batch, many measurements of
  uneven sizes in one persistent kernel per outer iteration (hipGraph replay)
* ``pnp_sci.unrolled``: unrolled_gap_tv / UnrolledGAPTV, the differentiable
  unrolled GAP-TV with hand-written backward kernels
"""

from .pnp_sci_algo import (ADMM_TV_rec, GAP_TV_rec, admm_denoise, admmdenoise_cacti,
                           gap_denoise, gap_denoise_bayer, solve_device)
from .tv import denoise_tv_chambolle
from .utils import A_, At_, make_operators, psnr

__all__ = ["gap_denoise", "admm_denoise", "admmdenoise_cacti", "GAP_TV_rec", "ADMM_TV_rec",
           "gap_denoise_bayer",
           "solve_device", "denoise_tv_chambolle", "A_", "At_", "psnr", "make_operators"]
