"""Drop-in ``wp``: the projected correlation function wp(rp)
(halotools/mock_observables/two_point_clustering/wp.py:21-240), from one
``rp_pi_tpcf`` with pi_bins = [0, pi_max]: wp = 2 pi_max xi(rp, pi < pi_max).
"""
from ._on_device import Residency
from .rp_pi_tpcf import _rp_pi_tpcf_process_args, rp_pi_tpcf

import numpy as np

__all__ = ("wp",)


def wp(sample1, rp_bins, pi_max, sample2=None, randoms=None, period=None, do_auto=True, do_cross=True,
       estimator="Natural", num_threads=1, approx_cell1_size=None, approx_cell2_size=None,
       approx_cellran_size=None, seed=None):
    """wp(rp), length len(rp_bins)-1, or the tuples the reference returns (wp.py:21-240).

    Accepts NumPy arrays (NumPy results) or CUDA tensors (CUDA tensor results).
    """
    pi_max = float(pi_max)
    pi_bins = np.array([0.0, pi_max])
    (sample1, rp_bins, pi_bins, sample2, randoms, period, do_auto, do_cross, num_threads, _sample1_is_sample2,
     PBCs) = _rp_pi_tpcf_process_args(sample1, rp_bins, pi_bins, sample2, randoms, period, do_auto, do_cross,
                                      estimator, num_threads, approx_cell1_size, approx_cell2_size,
                                      approx_cellran_size, seed)
    if _sample1_is_sample2:
        sample2 = None
    # the correlation functions stay on the device until the final scaling
    dev = Residency(sample1)
    result = rp_pi_tpcf(dev.put(sample1), rp_bins=rp_bins, pi_bins=pi_bins,
                        sample2=None if sample2 is None else dev.put(sample2),
                        randoms=dev.put(randoms), period=period, do_auto=do_auto, do_cross=do_cross,
                        estimator=estimator, num_threads=num_threads, approx_cell1_size=approx_cell1_size,
                        approx_cell2_size=approx_cell2_size, approx_cellran_size=approx_cellran_size)
    out = dev.out
    if _sample1_is_sample2:
        return out(2.0 * result[:, 0] * pi_max)
    if (do_auto is True) & (do_cross is True):
        return tuple(out(2.0 * r[:, 0] * pi_max) for r in result)
    elif (do_auto is True) & (do_cross is False):
        return tuple(out(2.0 * r[:, 0] * pi_max) for r in result)
    elif (do_auto is False) & (do_cross is True):
        return out(2.0 * result[:, 0] * pi_max)
