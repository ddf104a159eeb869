"""Drop-in ``tpcf``: the real-space two-point correlation function xi(r).

Same signature, checks, messages and results as the reference
(halotools/mock_observables/two_point_clustering/tpcf.py:212-601); the pair
counts come from the GPU ``npairs_3d`` and stay on the device through the
differencing and the estimator.
"""
from math import gamma
from warnings import warn

import numpy as np

from ...custom_exceptions import HalotoolsError
from ..mock_observables_helpers import (enforce_maximum_search_length, enforce_sample_has_correct_shape,
                                        get_num_threads, get_period, get_separation_bins_array)
from ..pair_counters import npairs_3d
from ._on_device import Residency, diff1
from .clustering_helpers import (process_optional_input_sample2, tpcf_estimator_dd_dr_rr_requirements,
                                 verify_tpcf_estimator)
from .tpcf_estimators import _TP_estimator, _TP_estimator_crossx

__all__ = ("tpcf",)


def _nball_volume(R, k=3):
    """Volume of a k-ball of radius R (tpcf.py:66-71)."""
    return (np.pi ** (k / 2.0) / gamma(k / 2.0 + 1.0)) * R ** k


def _random_counts_on(dev, sample1, sample2, randoms, rbins, period, PBCs, num_threads, do_RR, do_DR,
                      _sample1_is_sample2, approx_cell1_size, approx_cell2_size, approx_cellran_size, N1, N2):
    """tpcf.py:39-145 on device samples: counted randoms, or analytic ones (periodic box, no randoms)."""
    if randoms is not None:
        RR = D1R = D2R = None
        if do_RR is True:
            RR = diff1(npairs_3d(randoms, randoms, rbins, period=period, num_threads=num_threads,
                                 approx_cell1_size=approx_cellran_size, approx_cell2_size=approx_cellran_size))
        if do_DR is True:
            D1R = diff1(npairs_3d(sample1, randoms, rbins, period=period, num_threads=num_threads,
                                  approx_cell1_size=approx_cell1_size, approx_cell2_size=approx_cellran_size))
        if not _sample1_is_sample2 and do_DR is True:
            D2R = diff1(npairs_3d(sample2, randoms, rbins, period=period, num_threads=num_threads,
                                  approx_cell1_size=approx_cell2_size, approx_cell2_size=approx_cellran_size))
        return D1R, D2R, RR
    # analytic randoms: NR = N1 points, pair counts N * rho * shell volume
    NR = N1
    dv = np.diff(_nball_volume(rbins))
    global_volume = period.prod()
    rho1 = N1 / global_volume
    D1R = (NR) * (dv * rho1)
    rho2 = N2 / global_volume
    D2R = (NR) * (dv * rho2)
    rhor = (NR ** 2) / global_volume
    RR = dv * rhor
    return dev.put_f64(D1R), dev.put_f64(D2R), dev.put_f64(RR)


def _random_counts(sample1, sample2, randoms, rbins, period, PBCs, num_threads, do_RR, do_DR, _sample1_is_sample2,
                   approx_cell1_size, approx_cell2_size, approx_cellran_size):
    """tpcf.py:39-145 with the reference's arguments; results in the container of ``sample1``."""
    dev = Residency(sample1)
    s1 = dev.put(sample1)
    s2 = s1 if _sample1_is_sample2 else dev.put(sample2)
    res = _random_counts_on(dev, s1, s2, dev.put(randoms), rbins, period, PBCs, num_threads, do_RR, do_DR,
                            _sample1_is_sample2, approx_cell1_size, approx_cell2_size, approx_cellran_size,
                            len(sample1), len(sample2))
    return tuple(dev.out(r) for r in res)


def _pair_counts(sample1, sample2, rbins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                 approx_cell1_size, approx_cell2_size):
    """tpcf.py:148-209 with the reference's arguments; results in the container of ``sample1``."""
    dev = Residency(sample1)
    s1 = dev.put(sample1)
    s2 = s1 if _sample1_is_sample2 else dev.put(sample2)
    res = _pair_counts_on(s1, s2, rbins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                          approx_cell1_size, approx_cell2_size)
    return tuple(dev.out(r) for r in res)


def _pair_counts_on(sample1, sample2, rbins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                    approx_cell1_size, approx_cell2_size):
    """tpcf.py:148-209 on device samples (differences stay on the device)."""
    D1D1 = D1D2 = D2D2 = None
    if do_auto is True:
        D1D1 = diff1(npairs_3d(sample1, sample1, rbins, period=period, num_threads=num_threads,
                               approx_cell1_size=approx_cell1_size, approx_cell2_size=approx_cell1_size))
    if _sample1_is_sample2:
        D1D2 = D1D1
        D2D2 = D1D1
    else:
        if do_cross is True:
            D1D2 = diff1(npairs_3d(sample1, sample2, rbins, period=period, num_threads=num_threads,
                                   approx_cell1_size=approx_cell1_size, approx_cell2_size=approx_cell2_size))
        if do_auto is True:
            D2D2 = diff1(npairs_3d(sample2, sample2, rbins, period=period, num_threads=num_threads,
                                   approx_cell1_size=approx_cell2_size, approx_cell2_size=approx_cell2_size))
    return D1D1, D1D2, D2D2


def tpcf(sample1, rbins, sample2=None, randoms=None, period=None, do_auto=True, do_cross=True,
         estimator="Natural", num_threads=1, approx_cell1_size=None, approx_cell2_size=None,
         approx_cellran_size=None, RR_precomputed=None, NR_precomputed=None, seed=None):
    """xi(r) in the bins of ``rbins`` (length len(rbins)-1), or (xi_11, xi_12, xi_22),
    (xi_11, xi_22), xi_12 as the reference returns them (tpcf.py:212-498).

    Accepts NumPy arrays (NumPy results) or CUDA tensors (CUDA tensor results).
    ``num_threads``, ``approx_cell*_size`` and ``seed`` are validated as in the
    reference and do not change the result.
    """
    (sample1, rbins, sample2, randoms, period, do_auto, do_cross, num_threads, _sample1_is_sample2, PBCs,
     RR_precomputed, NR_precomputed) = _tpcf_process_args(
        sample1, rbins, sample2, randoms, period, do_auto, do_cross, estimator, num_threads, approx_cell1_size,
        approx_cell2_size, approx_cellran_size, RR_precomputed, NR_precomputed, seed)

    do_DD, do_DR, do_RR = tpcf_estimator_dd_dr_rr_requirements[estimator]
    if RR_precomputed is not None:
        do_RR = False

    N1 = len(sample1)
    N2 = len(sample2)
    if randoms is not None:
        NR = len(randoms)
    else:
        NR = NR_precomputed if NR_precomputed is not None else N1

    dev = Residency(sample1)
    s1 = dev.put(sample1)
    s2 = s1 if _sample1_is_sample2 else dev.put(sample2)
    rnd = dev.put(randoms)

    D1D1, D1D2, D2D2 = _pair_counts_on(s1, s2, rbins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                                       approx_cell1_size, approx_cell2_size)
    D1R, D2R, RR = _random_counts_on(dev, s1, s2, rnd, rbins, period, PBCs, num_threads, do_RR, do_DR,
                                     _sample1_is_sample2, approx_cell1_size, approx_cell2_size,
                                     approx_cellran_size, N1, N2)
    if RR_precomputed is not None:
        RR = dev.put_f64(RR_precomputed)

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


def _tpcf_process_args(sample1, rbins, sample2, randoms, period, do_auto, do_cross, estimator, num_threads,
                       approx_cell1_size, approx_cell2_size, approx_cellran_size, RR_precomputed, NR_precomputed,
                       seed):
    """tpcf.py:501-601, in the same order (so the same exception is raised first)."""
    sample1 = enforce_sample_has_correct_shape(sample1)
    sample2, _sample1_is_sample2, do_cross = process_optional_input_sample2(sample1, sample2, do_cross)
    if randoms is not None and not hasattr(randoms, "is_cuda"):
        randoms = np.atleast_1d(randoms)
    rbins = get_separation_bins_array(rbins)
    rmax = np.amax(rbins)
    period, PBCs = get_period(period)
    enforce_maximum_search_length(rmax, period)
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
    if (RR_precomputed is not None) | (NR_precomputed is not None):
        try:
            assert ((RR_precomputed is not None) & (NR_precomputed is not None)) is True
        except AssertionError:
            msg = ("\nYou must either provide both "
                   "``RR_precomputed`` and ``NR_precomputed`` arguments, or neither\n")
            raise HalotoolsError(msg)
        try:
            assert len(RR_precomputed) == len(rbins) - 1
        except AssertionError:
            msg = "\nLength of ``RR_precomputed`` must match length of ``rbins``\n"
            raise HalotoolsError(msg)
        if np.any(RR_precomputed == 0):
            msg = ("RR_precomputed has radial bin(s) which contain no pairs. \n"
                   "Consider increasing the number of randoms, or using larger bins.")
            warn(msg)
        try:
            assert len(randoms) == NR_precomputed
        except AssertionError:
            msg = ("If passing in randoms and also NR_precomputed, \n"
                   "the value of NR_precomputed must agree with the number of randoms\n")
            raise HalotoolsError(msg)
    assert np.all(rbins > 0.0), "All values of input ``rbins`` must be positive"
    return (sample1, rbins, sample2, randoms, period, do_auto, do_cross, num_threads, _sample1_is_sample2, PBCs,
            RR_precomputed, NR_precomputed)
