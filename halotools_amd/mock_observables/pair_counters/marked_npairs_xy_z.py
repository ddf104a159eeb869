"""Drop-in ``marked_npairs_xy_z`` on the MI355X pair-counting engine.

Same signature, checks, messages and output as the reference
(PC/marked_npairs_xy_z.py:23-333, PC = halotools/mock_observables/pair_counters).
The engine (PC/marked_cpairs/marked_npairs_xy_z_engine.pyx) and the marking
functions 1-15 it dispatches to (:231-263 there) are libhalopair.so.
"""
import numpy as np

from ... import _device as D
from ... import _engine as E
from ...custom_exceptions import HalotoolsError
from . import _common as C
from . import _process as P
from .marked_npairs_3d import _WFUNC_WIDTH, _marked_npairs_process_weights, same_weights

__all__ = ("marked_npairs_xy_z",)


def _func_signature_int_from_wfunc(weight_func_id):
    """PC/marked_npairs_xy_z.py:292-333: the xy_z engine knows ids 1-15 only."""
    if type(weight_func_id) != int:
        msg = "\n weight_func_id parameter must be an integer ID of a weighting function."
        raise ValueError(msg)
    if 1 <= weight_func_id <= 15:
        return _WFUNC_WIDTH[weight_func_id]
    msg = ("The value ``weight_func_id`` = %i is not recognized")
    raise HalotoolsError(msg % weight_func_id)


def marked_npairs_xy_z(sample1, sample2, rp_bins, pi_bins,
                       period=None, weights1=None, weights2=None,
                       weight_func_id=0, num_threads=1,
                       approx_cell1_size=None, approx_cell2_size=None):
    """Weighted sums W[k, g] of wfunc(w1_i, w2_j) over pairs with xy-separation
    <= ``rp_bins[k]`` and |dz| <= ``pi_bins[g]``.

    Returns float64 of shape ``(len(rp_bins), len(pi_bins))`` (NumPy, or a
    CUDA tensor when ``sample1`` is one), as the reference
    (PC/marked_npairs_xy_z.py:23-194).
    """
    # PC/marked_npairs_xy_z.py:150-160: the xy_z argument processing of npairs_xy_z
    num_threads = P.process_num_threads(num_threads)
    x1, y1, z1 = P.columns(sample1)
    x2, y2, z2 = P.columns(sample2)
    rp_bins = P.process_bins(rp_bins, P.RP_BINS_MSG)
    P.check_last_edge(rp_bins)
    rp_max = np.max(rp_bins)
    pi_bins = P.process_bins(pi_bins, P.PI_BINS_MSG)
    P.check_last_edge(pi_bins)
    pi_max = np.max(pi_bins)
    if period is None:
        pbc = False
        x1, y1, z1, x2, y2, z2, period = P.enclose_in_box(
            x1, y1, z1, x2, y2, z2, min_size=[rp_max * 3.0, rp_max * 3.0, pi_max * 3.0])
    else:
        pbc = True
        period = P.process_period(period)
    approx_cell1_size, approx_cell2_size = P.default_cell_sizes(
        approx_cell1_size, approx_cell2_size, (rp_max, rp_max, pi_max))
    xperiod, yperiod, zperiod = period
    # PC/marked_npairs_xy_z.py:161-164, then the mesh checks of :166-175
    raw_same = same_weights(weights1, weights2)
    weights1, weights2 = _marked_npairs_process_weights(sample1, sample2, weights1, weights2, weight_func_id,
                                                        signature=_func_signature_int_from_wfunc)
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rp_max, rp_max, pi_max), (xperiod, yperiod, zperiod))

    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    same = D.same_array(sample1, sample2) and (raw_same or same_weights(weights1, weights2))
    s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), D.same_array(sample1, sample2), device)
    nw = weights1.shape[1]
    w1 = D.to_device_weights(weights1, device)
    w2 = w1 if same else D.to_device_weights(weights2, device)
    mock = C.mock_dict(len(s1[0]), len(s2[0]), period, pbc)
    res = D.run(E.HP_MARKED_NPAIRS_XY_Z, [mock], s1, s2, P.squared(rp_bins), p2=P.squared(pi_bins), same=same,
                w1=w1, w2=w2, nw=nw, wfunc=weight_func_id, want_counts=False, want_sums=True, device=device)
    if as_tensor:
        return C.finish(res, lambda host: res.sums[0], True, period, pbc)
    return C.finish(res, lambda host: C.to_host(res.sums[0]), False, period, pbc)
