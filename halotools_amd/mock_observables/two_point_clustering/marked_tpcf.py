"""Drop-in ``marked_tpcf``: the marked correlation function M(r) = W(r) / R(r).

Same signature, checks, messages and results as the reference
(halotools/mock_observables/two_point_clustering/marked_tpcf.py:28-597); the
weighted and plain pair counts come from the GPU ``marked_npairs_3d`` /
``npairs_3d`` and stay on the device through the differencing and the ratio.

Two behaviours of the reference are kept on purpose, because its results
depend on them:
  * the shuffled marks of ``random_counts`` alias the marks they shuffle
    (marked_tpcf.py:420-428 assigns ``permuted_marks1 = marks1`` and permutes
    its columns in place), so the "random" weighted counts pair shuffled marks
    with the same shuffled marks, and with ``iterations > 1`` each iteration
    shuffles the previous iteration's marks again;
  * the permutations come from NumPy's global generator reseeded with
    ``seed`` inside an RNG context (``astropy.utils.misc.NumpyRNGContext``,
    restated in ``_numpy_rng_context``), so they equal the reference's draws.
"""
import contextlib

import numpy as np
import torch

from ...custom_exceptions import HalotoolsError
from ..mock_observables_helpers import (enforce_maximum_search_length, enforce_sample_has_correct_shape,
                                        get_num_threads, get_period, get_separation_bins_array)
from ..pair_counters import marked_npairs_3d, npairs_3d
from ._on_device import Residency, diff1
from .clustering_helpers import process_optional_input_sample2

__all__ = ("marked_tpcf",)


@contextlib.contextmanager
def _numpy_rng_context(seed):
    """astropy.utils.misc.NumpyRNGContext (astropy, absent here; its documented
    behaviour): seed NumPy's global generator with ``seed`` on entry (``None``
    reseeds from the OS) and restore the previous generator state on exit."""
    state = np.random.get_state()
    np.random.seed(seed)
    try:
        yield
    finally:
        np.random.set_state(state)


def _median0(x):
    """np.median(x, axis=0) on the device: middle value, or the mean of the two middle values."""
    s = torch.sort(x, dim=0).values
    n = s.shape[0]
    if n % 2:
        return s[n // 2]
    return (s[n // 2 - 1] + s[n // 2]) / 2.0


def marked_tpcf(sample1, rbins, sample2=None, marks1=None, marks2=None, period=None, do_auto=True, do_cross=True,
                num_threads=1, weight_func_id=1, normalize_by="random_marks", iterations=1, randomize_marks=None,
                seed=None):
    """Marked correlation function(s) in the bins of ``rbins`` (marked_tpcf.py:28-371).

    Accepts NumPy arrays (NumPy results) or CUDA tensors (CUDA tensor results).
    """
    (sample1, rbins, sample2, marks1, marks2, period, do_auto, do_cross, num_threads, weight_func_id, normalize_by,
     _sample1_is_sample2, PBCs, randomize_marks) = _marked_tpcf_process_args(
        sample1, rbins, sample2, marks1, marks2, period, do_auto, do_cross, num_threads, weight_func_id,
        normalize_by, iterations, randomize_marks, seed)

    dev = Residency(sample1)
    s1 = dev.put(sample1)
    s2 = s1 if _sample1_is_sample2 else dev.put(sample2)
    # the marks go to the device once (copies: the random counts permute them in place)
    m1 = dev.put_f64(marks1).clone()
    m2 = dev.put_f64(marks2).clone()

    W1W1, W1W2, W2W2 = marked_pair_counts(s1, s2, rbins, period, num_threads, do_auto, do_cross, m1, m2,
                                          weight_func_id, _sample1_is_sample2)
    if normalize_by == "number_counts":
        R1R1, R1R2, R2R2 = pair_counts(s1, s2, rbins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                                       None, None)
    elif iterations > 1:
        runs = []
        for i in range(iterations):
            with _numpy_rng_context(seed):
                permutate1 = np.random.permutation(np.arange(0, len(sample1)))
                permutate2 = np.random.permutation(np.arange(0, len(sample2)))
            runs.append(random_counts(s1, s2, rbins, period, num_threads, do_auto, do_cross, m1, m2, weight_func_id,
                                      _sample1_is_sample2, permutate1, permutate2, randomize_marks))
        # rows the reference fills with None hold NaN (NumPy's float conversion of None)
        nan = torch.full((len(rbins) - 1,), float("nan"), dtype=torch.float64, device=dev.device)
        R1R1, R1R2, R2R2 = [_median0(torch.stack([r[j] if r[j] is not None else nan for r in runs]))
                            for j in range(3)]
    else:
        with _numpy_rng_context(seed):
            permutate1 = np.random.permutation(np.arange(0, len(sample1)))
            permutate2 = np.random.permutation(np.arange(0, len(sample2)))
        R1R1, R1R2, R2R2 = random_counts(s1, s2, rbins, period, num_threads, do_auto, do_cross, m1, m2,
                                         weight_func_id, _sample1_is_sample2, permutate1, permutate2,
                                         randomize_marks)
    out = dev.out
    if _sample1_is_sample2:
        return out(W1W1 / R1R1)
    if (do_auto is True) & (do_cross is True):
        return out(W1W1 / R1R1), out(W1W2 / R1R2), out(W2W2 / R2R2)
    elif do_cross is True:
        return out(W1W2 / R1R2)
    elif do_auto is True:
        return out(W1W1 / R1R1), out(W2W2 / R2R2)


def marked_pair_counts(sample1, sample2, rbins, period, num_threads, do_auto, do_cross, marks1, marks2,
                       weight_func_id, _sample1_is_sample2):
    """marked_tpcf.py:374-408: weighted data pairs (device samples and marks)."""
    D1D1 = D1D2 = D2D2 = None
    if do_auto is True:
        D1D1 = diff1(marked_npairs_3d(sample1, sample1, rbins, weights1=marks1, weights2=marks1,
                                      weight_func_id=weight_func_id, period=period, num_threads=num_threads))
    if _sample1_is_sample2:
        D1D2 = D1D1
        D2D2 = D1D1
    else:
        if do_cross is True:
            D1D2 = diff1(marked_npairs_3d(sample1, sample2, rbins, weights1=marks1, weights2=marks2,
                                          weight_func_id=weight_func_id, period=period, num_threads=num_threads))
        if do_auto is True:
            D2D2 = diff1(marked_npairs_3d(sample2, sample2, rbins, weights1=marks2, weights2=marks2,
                                          weight_func_id=weight_func_id, period=period, num_threads=num_threads))
    return D1D1, D1D2, D2D2


def random_counts(sample1, sample2, rbins, period, num_threads, do_auto, do_cross, marks1, marks2, weight_func_id,
                  _sample1_is_sample2, permutate1, permutate2, randomize_marks):
    """marked_tpcf.py:411-456: weighted pairs with shuffled marks (the marks are
    shuffled in place; see the module docstring)."""
    permuted_marks1 = marks1
    permuted_marks2 = marks2
    p1 = torch.as_tensor(permutate1, device=marks1.device)
    p2 = torch.as_tensor(permutate2, device=marks2.device)
    for i in range(marks1.shape[1]):
        if randomize_marks[i]:
            permuted_marks1[:, i] = marks1[p1, i]
    for i in range(marks2.shape[1]):
        if randomize_marks[i]:
            permuted_marks2[:, i] = marks2[p2, i]
    R1R1 = R1R2 = R2R2 = None
    if do_auto is True:
        R1R1 = diff1(marked_npairs_3d(sample1, sample1, rbins, weights1=marks1, weights2=permuted_marks1,
                                      weight_func_id=weight_func_id, period=period, num_threads=num_threads))
    if _sample1_is_sample2:
        R1R2 = R1R1
        R2R2 = R1R1
    else:
        if do_cross is True:
            R1R2 = diff1(marked_npairs_3d(sample1, sample2, rbins, weights1=permuted_marks1, weights2=permuted_marks2,
                                          weight_func_id=weight_func_id, period=period, num_threads=num_threads))
        if do_auto is True:
            R2R2 = diff1(marked_npairs_3d(sample2, sample2, rbins, weights1=marks2, weights2=permuted_marks2,
                                          weight_func_id=weight_func_id, period=period, num_threads=num_threads))
    return R1R1, R1R2, R2R2


def pair_counts(sample1, sample2, rbins, period, num_threads, do_auto, do_cross, _sample1_is_sample2,
                approx_cell1_size, approx_cell2_size):
    """marked_tpcf.py:459-494: plain data pairs."""
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


def _as_marks(marks, n):
    if marks is None:
        return np.ones(n).astype(float)
    if isinstance(marks, torch.Tensor):
        return marks.to(torch.float64).clone()
    return np.atleast_1d(marks).astype(float)


def _marked_tpcf_process_args(sample1, rbins, sample2, marks1, marks2, period, do_auto, do_cross, num_threads,
                              wfunc, normalize_by, iterations, randomize_marks, seed):
    """marked_tpcf.py:497-597, in the same order."""
    sample1 = enforce_sample_has_correct_shape(sample1)
    sample2, _sample1_is_sample2, do_cross = process_optional_input_sample2(sample1, sample2, do_cross)
    try:
        int(wfunc) == wfunc
    except Exception:
        msg = ("\n `wfunc` parameter must be an integer ID of the desired function.")
        raise ValueError(msg)
    if normalize_by not in ["random_marks", "number_counts"]:
        msg = ("\n `normalize_by` parameter not recognized.")
        raise ValueError(msg)
    marks1 = _as_marks(marks1, len(sample1))
    marks2 = _as_marks(marks2, len(sample2))
    for k, marks in ((1, marks1), (2, marks2)):
        if marks.ndim not in (1, 2):
            msg = ("\n You must either pass in a 1-D or 2-D array \n"
                   "for the input `marks%d`. \n"
                   "The `pair_counters._wnpairs_process_weights` function received \n"
                   "a `marks%d` array of dimension %%i" % (k, k))
            raise HalotoolsError(msg % marks.ndim)
    if marks1.ndim == 1:
        marks1 = marks1.reshape((len(marks1), 1))
    if marks2.ndim == 1:
        marks2 = marks2.reshape((len(marks2), 1))
    if len(marks1) != len(sample1):
        msg = ("\n `marks1` must have same length as `sample1`.")
        raise HalotoolsError(msg)
    if len(marks2) != len(sample2):
        msg = ("\n `marks2` must have same length as `sample2`.")
        raise HalotoolsError(msg)
    if randomize_marks is not None:
        randomize_marks = np.atleast_1d(randomize_marks)
    else:
        randomize_marks = np.array([True] * marks1.shape[1])
    if randomize_marks.ndim == 1:
        if len(randomize_marks) != marks1.shape[1]:
            msg = ("\n `randomize_marks` must have same length \n"
                   " as the number of weights per point.")
            raise HalotoolsError(msg)
    else:
        msg = ("\n `randomize_marks` must be one dimensional.")
        raise HalotoolsError(msg)
    rbins = get_separation_bins_array(rbins)
    rmax = np.amax(rbins)
    period, PBCs = get_period(period)
    enforce_maximum_search_length(rmax, period)
    try:
        assert do_auto == bool(do_auto)
        assert do_cross == bool(do_cross)
    except Exception:
        msg = "`do_auto` and `do_cross` keywords must be boolean-valued."
        raise ValueError(msg)
    num_threads = get_num_threads(num_threads)
    return (sample1, rbins, sample2, marks1, marks2, period, do_auto, do_cross, num_threads, wfunc, normalize_by,
            _sample1_is_sample2, PBCs, randomize_marks)
