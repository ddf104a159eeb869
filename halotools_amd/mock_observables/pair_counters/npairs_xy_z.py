"""Drop-in ``npairs_xy_z`` on the MI355X pair-counting engine.

Same signature, checks, messages and output as the reference
(PC/npairs_xy_z.py:20-237, PC = halotools/mock_observables/pair_counters);
the engine (PC/cpairs/npairs_xy_z_engine.pyx) is libhalopair.so.
"""
import numpy as np

from ... import _device as D
from ... import _engine as E
from . import _common as C
from . import _process as P

__all__ = ("npairs_xy_z",)


def npairs_xy_z(sample1, sample2, rp_bins, pi_bins, period=None,
                num_threads=1, approx_cell1_size=None, approx_cell2_size=None):
    """Counts[k, g] of pairs with xy-separation <= ``rp_bins[k]`` and |dz| <= ``pi_bins[g]``.

    Returns int64 of shape ``(len(rp_bins), len(pi_bins))`` (NumPy, or a CUDA
    tensor when ``sample1`` is one), as the reference (PC/npairs_xy_z.py:20-163).
    """
    # PC/npairs_xy_z.py:166-237
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
    # PC/npairs_xy_z.py:128-145
    xperiod, yperiod, zperiod = period
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rp_max, rp_max, pi_max), (xperiod, yperiod, zperiod))

    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    same = D.same_array(sample1, sample2)
    s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), same, device)
    mock = C.mock_dict(len(s1[0]), len(s2[0]), period, pbc)
    res = D.run(E.HP_NPAIRS_XY_Z, [mock], s1, s2, P.squared(rp_bins), p2=P.squared(pi_bins), same=same,
                device=device)
    shape = (len(rp_bins), len(pi_bins))
    if as_tensor:
        return C.finish(res, lambda host: res.counts[0], True, period, pbc)
    return C.finish(res, lambda host: np.array(host[:-1].reshape(shape)), False, period, pbc)
