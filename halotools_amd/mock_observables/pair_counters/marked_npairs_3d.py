"""Drop-in ``marked_npairs_3d`` on the MI355X pair-counting engine.

Same signature, checks, messages and output as the reference
(PC/marked_npairs_3d.py:25-326, PC = halotools/mock_observables/pair_counters).
The engine (PC/marked_cpairs/marked_npairs_3d_engine.pyx) and the marking
functions (PC/marked_cpairs/marking_functions.pyx:14-217) are libhalopair.so.
"""
import numpy as np
import torch

from ... import _device as D
from ... import _engine as E
from ...custom_exceptions import HalotoolsError
from . import _common as C
from . import _process as P

__all__ = ("marked_npairs_3d",)

# number of weights per point of each weight_func_id (PC/marked_npairs_3d.py:290-323)
_WFUNC_WIDTH = {1: 1, 2: 1, 3: 2, 4: 2, 5: 2, 6: 2, 7: 2, 8: 2, 9: 2, 10: 2, 11: 2,
                12: 4, 13: 4, 14: 3, 15: 3, 16: 5, 17: 5}


def _func_signature_int_from_wfunc(weight_func_id):
    """PC/marked_npairs_3d.py:281-326."""
    if type(weight_func_id) != int:
        msg = "\n weight_func_id parameter must be an integer ID of a weighting function."
        raise ValueError(msg)
    if weight_func_id in _WFUNC_WIDTH:
        return _WFUNC_WIDTH[weight_func_id]
    msg = ("The value ``weight_func_id`` = %i is not recognized")
    raise HalotoolsError(msg % weight_func_id)


def _process_one(sample, weights, correct_shape, name, weight_func_id, correct_num_weights):
    npts_sample = correct_shape[0]
    converted = False
    if weights is None:
        return np.ones(correct_shape, dtype=np.float64)
    if D.is_device_tensor(weights):
        # device weights stay on the device (same checks and messages)
        weights = weights.to(torch.float64)
        weights = weights.reshape(1) if weights.ndim == 0 else weights
    else:
        weights = np.atleast_1d(weights).astype("float64")
    if weights.ndim == 1:
        converted = True
        weights = weights.reshape((len(weights), 1))
    elif weights.ndim != 2:
        msg = ("\n You must either pass in a 1-D or 2-D array \n"
               "for the input `%s`. Instead, an array of \n"
               "dimension %%i was received." % name)
        raise HalotoolsError(msg % weights.ndim)
    npts_weights, num_weights = tuple(weights.shape)[0], tuple(weights.shape)[1]
    if tuple(weights.shape) != correct_shape:
        if converted:
            msg = ("\n You passed in a 1-D array for `%s` that \n"
                   "does not have the correct length. The number of \n"
                   "points in `sample%s` = %%i, while the number of points \n"
                   "in your input 1-D `%s` array = %%i" % (name, name[-1], name))
            raise HalotoolsError(msg % (npts_sample, npts_weights))
        msg = ("\n You passed in a 2-D array for `%s` that \n"
               "does not have a consistent shape with `sample%s`. \n"
               "`sample%s` has length %%i. The input value of `weight_func_id` = %%i \n"
               "For this value of `weight_func_id`, there should be %%i weights \n"
               "per point. The shape of your input `%s` is (%%i, %%i)\n" % (name, name[-1], name[-1], name))
        raise HalotoolsError(msg % (npts_sample, weight_func_id, correct_num_weights, npts_weights, num_weights))
    return weights


def _marked_npairs_process_weights(sample1, sample2, weights1, weights2, weight_func_id,
                                   signature=_func_signature_int_from_wfunc):
    """PC/marked_npairs_3d.py:186-278 (PC/marked_npairs_xy_z.py:197-289 with its own `signature`)."""
    correct_num_weights = signature(weight_func_id)
    npts_sample1 = np.shape(sample1)[0]
    npts_sample2 = np.shape(sample2)[0]
    w1 = _process_one(sample1, weights1, (npts_sample1, correct_num_weights), "weights1", weight_func_id,
                      correct_num_weights)
    w2 = _process_one(sample2, weights2, (npts_sample2, correct_num_weights), "weights2", weight_func_id,
                      correct_num_weights)
    return w1, w2


def same_weights(w1, w2):
    """True when both weight arguments are known to hold the same values: the
    same object, the same view of one device tensor (device tensors are not
    compared by value, which would wait for the device), or equal host arrays.
    A False only costs the symmetric evaluation of an auto-count."""
    if w1 is w2:
        return True
    if isinstance(w1, torch.Tensor) or isinstance(w2, torch.Tensor):
        return D.same_array(w1, w2)
    try:
        return bool(np.array_equal(w1, w2))
    except (TypeError, ValueError):
        return False


def marked_npairs_3d(sample1, sample2, rbins, weight_func_id,
                     period=None, weights1=None, weights2=None, num_threads=1,
                     approx_cell1_size=None, approx_cell2_size=None):
    """Cumulative weighted pair sums W(<= rbins[k]) = sum of wfunc(w1_i, w2_j).

    ``weight_func_id`` selects one of the reference's marking functions 1-17
    (PC/marked_cpairs/marking_functions.pyx:14-217).  Returns float64 of length
    ``len(rbins)`` (NumPy), as the reference (PC/marked_npairs_3d.py:25-183).
    For differentiable product-weighted sums see
    ``halotools_amd.autodiff.marked_npairs_3d_product``.
    """
    # PC/marked_npairs_3d.py:141-183 (the 3-D argument processing is shared with npairs_3d)
    num_threads = P.process_num_threads(num_threads)
    x1, y1, z1 = P.columns(sample1)
    x2, y2, z2 = P.columns(sample2)
    rbins = np.atleast_1d(rbins).astype("f8")
    rmax = np.max(rbins)
    rbins = P.process_bins(rbins, P.RBINS_MSG)
    P.check_last_edge(rbins)
    if period is None:
        pbc = False
        x1, y1, z1, x2, y2, z2, period = P.enclose_in_box(
            x1, y1, z1, x2, y2, z2, min_size=[rmax * 3.0, rmax * 3.0, rmax * 3.0])
    else:
        pbc = True
        period = P.process_period(period)
    approx_cell1_size, approx_cell2_size = P.default_cell_sizes(
        approx_cell1_size, approx_cell2_size, (rmax, rmax, rmax))
    xperiod, yperiod, zperiod = period
    raw_same = same_weights(weights1, weights2)
    weights1, weights2 = _marked_npairs_process_weights(sample1, sample2, weights1, weights2, weight_func_id)
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rmax, rmax, rmax), (xperiod, yperiod, zperiod))

    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    same = D.same_array(sample1, sample2) and (raw_same or same_weights(weights1, weights2))
    s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), D.same_array(sample1, sample2), device)
    nw = weights1.shape[1]
    w1 = D.to_device_weights(weights1, device)
    w2 = w1 if same else D.to_device_weights(weights2, device)
    mock = C.mock_dict(len(s1[0]), len(s2[0]), period, pbc)
    res = D.run(E.HP_MARKED_NPAIRS_3D, [mock], s1, s2, P.squared(rbins), same=same, w1=w1, w2=w2, nw=nw,
                wfunc=weight_func_id, want_counts=False, want_sums=True, device=device)
    if as_tensor:
        return C.finish(res, lambda host: res.sums[0], True, period, pbc)
    return C.finish(res, lambda host: C.to_host(res.sums[0]), False, period, pbc)
