"""Drop-in ``npairs_3d`` on the MI355X pair-counting engine.

Same signature, argument checks, error messages and output as the reference
``halotools.mock_observables.npairs_3d`` (PC/npairs_3d.py:20-213, PC =
halotools/mock_observables/pair_counters).  The RectangularDoubleMesh build
and the Cython engine (PC/cpairs/npairs_3d_engine.pyx) are replaced by
libhalopair.so: a GPU counting-sort cell build and the FP64 pair kernel.
"""
import numpy as np

from ... import _device as D
from ... import _engine as E
from . import _common as C
from . import _process as P

__all__ = ("npairs_3d",)


def npairs_3d(sample1, sample2, rbins, period=None,
              num_threads=1, approx_cell1_size=None, approx_cell2_size=None):
    """Cumulative number of ordered pairs with 3-D separation <= each ``rbins`` edge.

    Parameters and return value as in the reference (PC/npairs_3d.py:20-109):
    ``sample1``/``sample2`` are (N, 3) arrays, ``rbins`` the edges, ``period``
    None (non-periodic) or a scalar / 3-vector; ``num_threads`` and the
    ``approx_cell*_size`` hints are validated like the reference and otherwise
    do not change the result.  Auto-counts include self-pairs and count each
    pair twice, exactly like the reference.

    Returns an int64 NumPy array of length ``len(rbins)``; when ``sample1`` is a
    CUDA tensor, an int64 CUDA tensor on its device instead.
    """
    # PC/npairs_3d.py:153-213
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
    # PC/npairs_3d.py:116-132
    xperiod, yperiod, zperiod = period
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rmax, rmax, rmax), (xperiod, yperiod, zperiod))

    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    same = D.same_array(sample1, sample2)
    s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), same, device)
    mock = C.mock_dict(len(s1[0]), len(s2[0]), period, pbc)
    res = D.run(E.HP_NPAIRS_3D, [mock], s1, s2, P.squared(rbins), same=same, device=device)
    if as_tensor:
        return C.finish(res, lambda host: res.counts[0], True, period, pbc)
    return C.finish(res, lambda host: np.array(host[:-1].reshape(len(rbins))), False, period, pbc)
