"""Drop-in ``npairs_per_object_3d`` on the MI355X pair-counting engine.

Same signature, argument checks, error messages and output as the reference
(PC/npairs_per_object_3d.py:20-137, PC = halotools/mock_observables/pair_counters),
which shares its argument processing with npairs_3d (PC/npairs_3d.py:153-213).
The RectangularDoubleMesh build, the Cython engine
(PC/cpairs/npairs_per_object_3d_engine.pyx:19-208) and its unsorting are
libhalopair.so: each 64-point lane group keeps one count row per lane point and
writes it, cumulative over the bins, at the point's input position.
"""
import numpy as np

from ... import _device as D
from ... import _engine as E
from . import _common as C
from . import _process as P

__all__ = ("npairs_per_object_3d",)


def npairs_per_object_3d(sample1, sample2, rbins, period=None,
                         num_threads=1, approx_cell1_size=None, approx_cell2_size=None):
    """Counts of ``sample2`` points within each ``rbins`` edge of every ``sample1`` point.

    Parameters as in npairs_3d (PC/npairs_per_object_3d.py:20-80).  Returns an
    int64 array of shape (Npts1, len(rbins)) in the order of ``sample1`` (NumPy,
    or a CUDA tensor on the input's device when ``sample1`` is one).
    """
    # PC/npairs_3d.py:153-213 (shared argument processing)
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
    # PC/npairs_per_object_3d.py:103-118
    xperiod, yperiod, zperiod = period
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rmax, rmax, rmax), (xperiod, yperiod, zperiod))

    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    same = D.same_array(sample1, sample2)
    s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), same, device)
    n1 = len(s1[0])
    mock = C.mock_dict(n1, len(s2[0]), period, pbc)
    res = D.run(E.HP_NPAIRS_PER_OBJECT_3D, [mock], s1, s2, P.squared(rbins), same=same, device=device)
    if as_tensor:
        return C.finish(res, lambda host: res.counts, True, period, pbc)
    return C.finish(res, lambda host: np.array(host[:-1].reshape(n1, len(rbins))), False, period, pbc)
