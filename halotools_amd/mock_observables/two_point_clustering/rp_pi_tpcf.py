"""Drop-in ``rp_pi_tpcf``: the redshift-space correlation function xi(rp, pi).

Same signature, checks, messages and results as the reference
(halotools/mock_observables/two_point_clustering/rp_pi_tpcf.py:23-535); the
pair counts come from the GPU ``npairs_xy_z`` and stay on the device through
the 2-D differencing and the estimator.
"""
from math import pi

import numpy as np

from ..mock_observables_helpers import (enforce_maximum_search_length, enforce_sample_has_correct_shape,
                                        get_line_of_sight_bins_array, get_num_threads, get_period,
                                        get_separation_bins_array)
from ..pair_counters import npairs_xy_z
from ._on_device import Residency, diff2
from .clustering_helpers import process_optional_input_sample2, verify_tpcf_estimator
from .tpcf_estimators import _TP_estimator, _TP_estimator_crossx, _TP_estimator_requirements

__all__ = ("rp_pi_tpcf",)


def rp_pi_tpcf(sample1, rp_bins, pi_bins, sample2=None, randoms=None, period=None, do_auto=True, do_cross=True,
               estimator="Natural", num_threads=1, approx_cell1_size=None, approx_cell2_size=None,
               approx_cellran_size=None, seed=None):
    """xi(rp, pi), shape (len(rp_bins)-1, len(pi_bins)-1), or the tuples the
    reference returns for cross-correlations (rp_pi_tpcf.py:23-288).

    Accepts NumPy arrays (NumPy results) or CUDA tensors (CUDA tensor results).
    """
    (sample1, rp_bins, pi_bins, sample2, randoms, period, do_auto, do_cross, num_threads, _sample1_is_sample2,
     PBCs) = _rp_pi_tpcf_process_args(sample1, rp_bins, pi_bins, sample2, randoms, period, do_auto, do_cross,
                                      estimator, num_threads, approx_cell1_size, approx_cell2_size,
                                      approx_cellran_size, seed)
    do_DD, do_DR, do_RR = _TP_estimator_requirements(estimator)
    N1 = len(sample1)
    N2 = len(sample2)
    NR = len(randoms) if randoms is not None else N1

    dev = Residency(sample1)
    s1 = dev.put(sample1)
    s2 = s1 if _sample1_is_sample2 else dev.put(sample2)
    rnd = dev.put(randoms)
    D1D1, D1D2, D2D2 = _pair_counts_on(s1, s2, rp_bins, pi_bins, period, num_threads, do_auto, do_cross,
                                       _sample1_is_sample2, approx_cell1_size, approx_cell2_size)
    D1R, D2R, RR = _random_counts_on(dev, s1, s2, rnd, rp_bins, pi_bins, period, PBCs, num_threads, do_RR, do_DR,
                                     _sample1_is_sample2, approx_cell1_size, approx_cell2_size,
                                     approx_cellran_size, N1, N2)
    out = dev.out
    if _sample1_is_sample2:
        return out(_TP_estimator(D1D1, D1R, RR, N1, N1, NR, NR, estimator))
    if (do_auto is True) & (do_cross is True):
        xi_11 = _TP_estimator(D1D1, D1R, RR, N1, N1, NR, NR, estimator)
        xi_12 = _TP_estimator_crossx(D1D2, D1R, D2R, RR, N1, N2, NR, NR, estimator)
        xi_22 = _TP_estimator(D2D2, D2R, RR, N2, N2, NR, NR, estimator)
        return out(xi_11), out(xi_12), out(xi_22)
    elif do_cross is True:
        return out(_TP_estimator_crossx(D1D2, D1R, D2R, RR, N1, N2, NR, NR, estimator))
    elif do_auto is True:
        xi_11 = _TP_estimator(D1D1, D1R, RR, N1, N1, NR, NR, estimator)
        xi_22 = _TP_estimator(D2D2, D2R, RR, N2, N2, NR, NR, estimator)
        return out(xi_11), out(xi_22)


def _pair_counts_on(sample1, sample2, rp_bins, pi_bins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                    approx_cell1_size, approx_cell2_size):
    """rp_pi_tpcf.py:291-350 (D1D1 is always counted there, whatever ``do_auto``)."""
    D1D1 = diff2(npairs_xy_z(sample1, sample1, rp_bins, pi_bins, period=period, num_threads=num_threads,
                             approx_cell1_size=approx_cell1_size, approx_cell2_size=approx_cell1_size))
    D1D2 = D2D2 = None
    if _sample1_is_sample2:
        D1D2 = D1D1
        D2D2 = D1D1
    else:
        if do_cross is True:
            D1D2 = diff2(npairs_xy_z(sample1, sample2, rp_bins, pi_bins, period=period, num_threads=num_threads,
                                     approx_cell1_size=approx_cell1_size, approx_cell2_size=approx_cell2_size))
        if do_auto is True:
            D2D2 = diff2(npairs_xy_z(sample2, sample2, rp_bins, pi_bins, period=period, num_threads=num_threads,
                                     approx_cell1_size=approx_cell2_size, approx_cell2_size=approx_cell2_size))
    return D1D1, D1D2, D2D2


def cylinder_volume(R, h):
    """Volumes pi R^2 h of cylinders, outer product over (R, h) (rp_pi_tpcf.py:353-357)."""
    return pi * np.outer(R ** 2.0, h)


def _random_counts_on(dev, sample1, sample2, randoms, rp_bins, pi_bins, period, PBCs, num_threads, do_RR, do_DR,
                      _sample1_is_sample2, approx_cell1_size, approx_cell2_size, approx_cellran_size, N1, N2):
    """rp_pi_tpcf.py:360-460: counted randoms, or analytic ones (periodic box, no randoms)."""
    if randoms is not None:
        RR = D1R = D2R = None
        if do_RR is True:
            RR = diff2(npairs_xy_z(randoms, randoms, rp_bins, pi_bins, period=period, num_threads=num_threads,
                                   approx_cell1_size=approx_cellran_size, approx_cell2_size=approx_cellran_size))
        if do_DR is True:
            D1R = diff2(npairs_xy_z(sample1, randoms, rp_bins, pi_bins, period=period, num_threads=num_threads,
                                    approx_cell1_size=approx_cell1_size, approx_cell2_size=approx_cellran_size))
        if not _sample1_is_sample2 and do_DR is True:
            D2R = diff2(npairs_xy_z(sample2, randoms, rp_bins, pi_bins, period=period, num_threads=num_threads,
                                    approx_cell1_size=approx_cell2_size, approx_cell2_size=approx_cellran_size))
        return D1R, D2R, RR
    NR = N1
    v = cylinder_volume(rp_bins, 2.0 * pi_bins)
    dv = np.diff(np.diff(v, axis=0), axis=1)  # volumes of the (rp, pi) annuli
    global_volume = period.prod()
    rho1 = N1 / global_volume
    D1R = (N1) * (dv * rho1)
    rho2 = N2 / global_volume
    D2R = N2 * (dv * rho2)
    rhor = NR ** 2 / global_volume
    RR = dv * rhor
    return dev.put_f64(D1R), dev.put_f64(D2R), dev.put_f64(RR)


def _rp_pi_tpcf_process_args(sample1, rp_bins, pi_bins, sample2, randoms, period, do_auto, do_cross, estimator,
                             num_threads, approx_cell1_size, approx_cell2_size, approx_cellran_size, seed):
    """rp_pi_tpcf.py:463-535, in the same order."""
    sample1 = enforce_sample_has_correct_shape(sample1)
    sample2, _sample1_is_sample2, do_cross = process_optional_input_sample2(sample1, sample2, do_cross)
    if randoms is not None and not hasattr(randoms, "is_cuda"):
        randoms = np.atleast_1d(randoms)
    rp_bins = get_separation_bins_array(rp_bins)
    rp_max = np.amax(rp_bins)
    pi_bins = get_line_of_sight_bins_array(pi_bins)
    pi_max = np.amax(pi_bins)
    period, PBCs = get_period(period)
    enforce_maximum_search_length([rp_max, rp_max, pi_max], period)
    if (randoms is None) & (PBCs is False):
        msg = "If no PBCs are specified, randoms must be provided.\n"
        raise ValueError(msg)
    try:
        assert do_auto == bool(do_auto)
        assert do_cross == bool(do_cross)
    except Exception:
        msg = "`do_auto` and `do_cross` keywords must be boolean-valued."
        raise ValueError(msg)
    num_threads = get_num_threads(num_threads)
    verify_tpcf_estimator(estimator)
    return (sample1, rp_bins, pi_bins, sample2, randoms, period, do_auto, do_cross, num_threads,
            _sample1_is_sample2, PBCs)
