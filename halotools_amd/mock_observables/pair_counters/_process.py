"""Argument processing of the pair counters, restated from the reference.

Every check and message below follows the reference so that a caller who
switches from ``halotools.mock_observables`` sees the same exceptions in the
same order (PC/ = halotools/mock_observables/pair_counters/):

* ``process_num_threads``      -- PC/npairs_3d.py:157-162
* ``process_bins``             -- PC/npairs_3d.py:171-182, PC/npairs_xy_z.py:183-206
* ``process_period``           -- PC/npairs_3d.py:185-200 (PBC branch)
* ``enclose_in_box``           -- PC/mesh_helpers.py:17-64 (non-PBC branch)
* ``default_cell_sizes``       -- PC/npairs_3d.py:202-209
* ``set_approximate_cell_sizes`` -- PC/mesh_helpers.py:113-145
* ``check_search_lengths``     -- PC/rectangular_mesh.py:376-417
* ``array_is_monotonic`` / ``custom_len`` -- halotools/utils/array_utils.py:14-53,135-186

The cell sizes are validated exactly as the reference validates them and are
then only hints: the GPU mesh is independent of them, and so are the counts.
"""
import multiprocessing

import numpy as np
import torch

from ...custom_exceptions import HalotoolsError


def custom_len(x):
    """halotools/utils/array_utils.py:14-53."""
    try:
        return len(x)
    except TypeError:
        return 1


def array_is_monotonic(array, strict=False):
    """halotools/utils/array_utils.py:135-186."""
    if custom_len(array) < 3:
        raise HalotoolsError("Input array to the array_is_monotonic method has less then 3 elements")
    if strict is True and isinstance(array, np.ndarray) and array.ndim == 1:
        # the common case of a checked bin array, in one comparison pass
        if bool((array[1:] > array[:-1]).all()):
            return 1
    d = np.diff(array)
    if strict is True:
        if np.all(d > 0):
            return 1
        elif np.all(d < 0):
            return -1
        return 0
    if np.all(d >= 0):
        return 1
    elif np.all(d <= 0):
        return -1
    return 0


def process_num_threads(num_threads):
    """PC/npairs_3d.py:157-162 (the value is validated, then unused on the GPU)."""
    # the reference tests ``num_threads is not 1``: only the int 1 itself skips the checks
    if not (type(num_threads) is int and num_threads == 1):
        if num_threads == "max":
            num_threads = multiprocessing.cpu_count()
        if not isinstance(num_threads, int):
            msg = "Input ``num_threads`` argument must be an integer or the string 'max'"
            raise ValueError(msg)
    return num_threads


def process_bins(bins, msg):
    """atleast_1d(...).astype('f8'), >= 2 entries, strictly increasing when > 2."""
    bins = np.atleast_1d(bins).astype("f8")
    try:
        assert bins.ndim == 1
        assert len(bins) > 1
        if len(bins) > 2:
            assert array_is_monotonic(bins, strict=True) == 1
    except AssertionError:
        raise ValueError(msg)
    return bins


def check_last_edge(bins):
    """Reject the one bin layout whose counts depend on the reference's mesh.

    The reference only checks monotonicity for more than two edges
    (PC/npairs_3d.py:178-179) and sizes its search by max(bins); a pair is
    counted in the last bin when dsq <= bins[-1]**2.  If bins[-1]**2 exceeds
    max(bins)**2 (a descending two-edge array with a negative edge), pairs
    beyond the search length would count only where its mesh happens to
    reach, so the result is not defined by the inputs.
    """
    m = bins.max()
    if bins[-1] * bins[-1] > m * m:
        raise ValueError("Input bins whose last edge squared exceeds max(bins)**2 are not supported: "
                         "the reference's result for them depends on its mesh")


RBINS_MSG = "Input ``rbins`` must be a monotonically increasing 1D array with at least two entries"
RP_BINS_MSG = ("Input ``rp_bins`` must be a monotonically increasing 1D array "
               "with at least two entries")
PI_BINS_MSG = "Input ``pi_bins`` must be a monotonically increasing 1D array with at least two entries"


def process_period(period):
    """PC/npairs_3d.py:190-200."""
    if type(period) in (float, int):
        # a Python scalar: the same array and checks without NumPy's
        # conversions of a one-element array
        p = float(period)
        if not (p < np.inf and p > 0):
            raise ValueError("Input ``period`` must be a bounded positive number in all dimensions")
        return np.array([p, p, p])
    period = np.atleast_1d(period).astype(float)
    if len(period) == 1:
        period = np.array([period[0]] * 3)
    try:
        assert np.all(period < np.inf)
        assert np.all(period > 0)
    except AssertionError:
        msg = "Input ``period`` must be a bounded positive number in all dimensions"
        raise ValueError(msg)
    return period


def _min(a):
    return a.min() if isinstance(a, torch.Tensor) else np.min(a)


def _max(a):
    return a.max() if isinstance(a, torch.Tensor) else np.max(a)


def enclose_in_box(x1, y1, z1, x2, y2, z2, min_size=None):
    """PC/mesh_helpers.py:17-64: shift all six arrays by the one scalar minimum.

    The subtraction runs in the input dtype (NumPy on the host or torch on the
    device, identical IEEE operations) before anything is converted to float64,
    which is what the reference does.
    """
    if isinstance(x1, torch.Tensor):
        arrs = (x1, x2, y1, y2, z1, z2)
        xyzmin = torch.stack([_min(a) for a in arrs]).min()
        # max - min in the input dtype, as the reference computes it
        xyzmax = float((torch.stack([_max(a) for a in arrs]).max() - xyzmin).item())
    else:
        xyzmin = np.min([np.min(x1), np.min(x2), np.min(y1), np.min(y2), np.min(z1), np.min(z2)])
        xyzmax = np.max([np.max(x1), np.max(x2), np.max(y1), np.max(y2), np.max(z1), np.max(z2)]) - xyzmin
    x1, y1, z1 = x1 - xyzmin, y1 - xyzmin, z1 - xyzmin
    x2, y2, z2 = x2 - xyzmin, y2 - xyzmin, z2 - xyzmin
    Lbox = np.array([xyzmax, xyzmax, xyzmax], dtype=float)
    if min_size is not None:
        min_size = np.atleast_1d(min_size)
        if np.any(Lbox < min_size):
            Lbox[(Lbox < min_size)] = min_size[(Lbox < min_size)]
    return x1, y1, z1, x2, y2, z2, Lbox


def default_cell_sizes(approx_cell1_size, approx_cell2_size, default):
    """PC/npairs_3d.py:202-209."""
    if approx_cell1_size is None:
        approx_cell1_size = list(default)
    elif custom_len(approx_cell1_size) == 1:
        approx_cell1_size = [approx_cell1_size, approx_cell1_size, approx_cell1_size]
    if approx_cell2_size is None:
        approx_cell2_size = list(default)
    elif custom_len(approx_cell2_size) == 1:
        approx_cell2_size = [approx_cell2_size, approx_cell2_size, approx_cell2_size]
    return approx_cell1_size, approx_cell2_size


def set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period):
    """PC/mesh_helpers.py:113-145."""
    if approx_cell1_size is None:
        approx_cell1_size = period / 10.0
    else:
        approx_cell1_size = np.atleast_1d(approx_cell1_size)
        try:
            assert len(approx_cell1_size) == 3
            assert type(approx_cell1_size) is np.ndarray
            assert approx_cell1_size.ndim == 1
        except AssertionError:
            raise ValueError("Input ``approx_cell1_size`` must be a length-3 sequence")
    if approx_cell2_size is None:
        approx_cell2_size = np.copy(approx_cell1_size)
    else:
        approx_cell2_size = np.atleast_1d(approx_cell2_size)
        try:
            assert len(approx_cell2_size) == 3
            assert type(approx_cell2_size) is np.ndarray
            assert approx_cell2_size.ndim == 1
        except AssertionError:
            raise ValueError("Input ``approx_cell2_size`` must be a length-3 sequence")
    return approx_cell1_size, approx_cell2_size


def check_search_lengths(search, period):
    """PC/rectangular_mesh.py:376-417 (RectangularDoubleMesh._check_sensible_constructor_inputs)."""
    for name, s, p in zip("xyz", search, period):
        if not s <= p / 3.0:
            msg = ("\n The maximum length over which you search for pairs of points \n"
                   "cannot be larger than Lbox/3 in any dimension. \n"
                   "You tried to search for pairs out to a length of search_%slength = %.2f,\n"
                   "but the size of your box in this dimension is %speriod = %.2f.\n"
                   "If you need to count pairs on these length scales, \n"
                   "you should use a larger simulation.\n" % (name, s, name, p))
            raise ValueError(msg)


def columns(sample):
    """``sample[:, 0]`` ... as the reference reads them (PC/npairs_3d.py:165-170)."""
    if isinstance(sample, torch.Tensor) and sample.dim() == 2 and sample.shape[1] == 3:
        return sample.unbind(1)  # the same three column views in one call
    return sample[:, 0], sample[:, 1], sample[:, 2]


def squared(bins):
    """rbins*rbins exactly as the engines form it (PC/cpairs/npairs_3d_engine.pyx:46)."""
    return bins * bins
