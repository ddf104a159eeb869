"""Argument processing shared by the two-point functions, restated from the
reference (halotools/mock_observables/mock_observables_helpers.py:74-206 and
PC/mesh_helpers.py:224-253, PC = halotools/mock_observables/pair_counters).

Samples may be NumPy arrays (or anything ``np.atleast_1d`` accepts) or CUDA
tensors; the checks and messages are the reference's either way.
"""
import multiprocessing
from warnings import warn

import numpy as np
import torch

from .pair_counters._process import array_is_monotonic

num_available_cores = multiprocessing.cpu_count()

__all__ = ("get_num_threads", "get_period", "enforce_sample_has_correct_shape", "get_separation_bins_array",
           "get_line_of_sight_bins_array")


def is_device(a):
    return isinstance(a, torch.Tensor) and a.is_cuda


def get_num_threads(input_num_threads, enforce_max_cores=False):
    """mock_observables_helpers.py:74-104 (validated, then unused on the GPU)."""
    if input_num_threads == "max":
        num_threads = num_available_cores
    else:
        try:
            num_threads = int(input_num_threads)
            assert num_threads == input_num_threads
        except Exception:
            msg = "Input ``num_threads`` must be an integer"
            raise ValueError(msg)
    if num_threads > num_available_cores:
        msg = "Input ``num_threads`` = {0} exceeds the ``num_available_cores`` = {1}.\n"
        warn(msg.format(num_threads, num_available_cores))
        if enforce_max_cores is True:
            msg = ("Since ``enforce_max_cores`` is True,\n"
                   "setting ``num_threads`` to ``num_available_cores``.")
            warn(msg)
            num_threads = num_available_cores
    return num_threads


def get_period(period):
    """mock_observables_helpers.py:107-132: (None, False) or (float[3], True)."""
    if period is None:
        PBCs = False
    else:
        PBCs = True
        period = np.atleast_1d(period).astype(float)
        if len(period) == 1:
            period = np.array([period[0]] * 3).astype(float)
        try:
            assert np.all(period < np.inf)
            assert np.all(period > 0)
            assert len(period) == 3
        except AssertionError:
            msg = ("Input ``period`` must be either a scalar or a 3-element sequence.\n"
                   "All values must bounded positive numbers.\n")
            raise ValueError(msg)
    return period, PBCs


def enforce_sample_has_correct_shape(sample, ndim=3):
    """mock_observables_helpers.py:135-150: shape (Npts, ndim) or TypeError."""
    if not isinstance(sample, torch.Tensor):
        sample = np.atleast_1d(sample)
    try:
        input_shape = tuple(sample.shape)
        assert len(input_shape) == 2
        assert input_shape[1] == ndim
    except AssertionError:
        msg = ("Input sample of points must be a Numpy ndarray of shape (Npts, {0}).\n"
               "To convert a sequence of 1d arrays x, y, z into correct shape expected \n"
               "throughout the `mock_observables` package:\n\n"
               ">>> sample = np.vstack([x, y, z]).T ".format(ndim))
        raise TypeError(msg)
    return sample


def get_separation_bins_array(separation_bins):
    """mock_observables_helpers.py:153-177: >= 2 strictly increasing positive edges."""
    separation_bins = np.atleast_1d(separation_bins)
    try:
        assert separation_bins.ndim == 1
        assert len(separation_bins) > 1
        if len(separation_bins) > 2:
            assert array_is_monotonic(separation_bins, strict=True) == 1
        assert np.all(separation_bins > 0)
    except AssertionError:
        msg = ("\n Input separation bins must be a monotonically increasing \n"
               "1-D array with at least two entries, all of which must be strictly positive.\n")
        raise TypeError(msg)
    return separation_bins


def get_line_of_sight_bins_array(pi_bins):
    """mock_observables_helpers.py:180-206: as above, zero allowed."""
    pi_bins = np.atleast_1d(pi_bins)
    try:
        assert pi_bins.ndim == 1
        assert len(pi_bins) > 1
        if len(pi_bins) > 2:
            assert array_is_monotonic(pi_bins, strict=True) == 1
    except AssertionError:
        msg = ("\n Input separation bins must be a monotonically increasing \n"
               "1-D array with at least two entries.\n")
        raise TypeError(msg)
    return pi_bins


def enforce_maximum_search_length(search_length, period=None):
    """PC/mesh_helpers.py:224-253: every search length strictly below period/3."""
    search_length = np.atleast_1d(search_length)
    if period is None:
        period = np.zeros_like(search_length) + np.inf
    period = np.atleast_1d(period)
    try:
        assert np.all(search_length < period / 3.)
    except AssertionError:
        max_search_fraction = np.max(period - search_length)
        msg = ("The search algorithm used by the function you called \n"
               "does not permit you to look for pairs separated by values \n"
               "exceeding a fraction of Lbox/3. in any dimension.\n"
               "Your function call would require searching for pairs separated by a distance of {0:.2f}*Lbox.\n"
               "Either decrease your search length or use a larger simulation.")
        raise ValueError(msg.format(max_search_fraction))
