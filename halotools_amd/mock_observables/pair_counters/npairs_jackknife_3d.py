"""Drop-in ``npairs_jackknife_3d`` and ``npairs_jackknife_xy_z`` on the MI355X
pair-counting engine.

Same signatures, checks, messages and outputs as the reference
(PC/npairs_jackknife_3d.py:24-276 and PC/npairs_jackknife_xy_z.py:23-276,
PC = halotools/mock_observables/pair_counters); the engines
(PC/cpairs/npairs_jackknife_3d_engine.pyx, npairs_jackknife_xy_z_engine.pyx)
are libhalopair.so.  Row s of the result is what the reference's engine sums
with its jackknife weight for subsample s (row 0: every pair, w1*w2).
"""
from warnings import warn

import numpy as np
import torch

from ... import _device as D
from ... import _engine as E
from ...custom_exceptions import HalotoolsError
from . import _common as C
from . import _process as P

__all__ = ("npairs_jackknife_3d", "npairs_jackknife_xy_z")


def _n(a):
    return int(tuple(a.shape)[0]) if hasattr(a, "shape") else int(np.shape(a)[0])


def _as_weights(w, sample, name):
    if w is None:
        return None
    if isinstance(w, torch.Tensor):
        w = w.to(torch.float64)
    else:
        w = np.asarray(w).astype("float64")
    if _n(w) != _n(sample):
        raise HalotoolsError("%s should have same len as %s" % (name, "sample" + name[-1]))
    if w.ndim != 1:
        # the reference's engine takes one weight per point (a 1-D buffer) and
        # rejects other shapes when it casts them
        raise ValueError("Buffer has wrong number of dimensions (expected 1, got %d)" % w.ndim)
    return w


def _as_tags(t, sample, name):
    t = t.to(torch.int64) if isinstance(t, torch.Tensor) else np.asarray(t).astype("int")
    if _n(t) != _n(sample):
        raise HalotoolsError("%s should have same len as %s" % (name, "sample" + name[-1]))
    return t


def _min(t):
    return int(t.min()) if isinstance(t, torch.Tensor) else int(np.min(t))


def _max(t):
    return int(t.max()) if isinstance(t, torch.Tensor) else int(np.max(t))


def _covers_all(t, N_samples):
    """np.array_equal(np.unique(t), np.arange(1, N_samples + 1)) for tags
    already checked to lie in [1, N_samples]: every tag value occurs.  A
    histogram instead of a sort (no more bins than points: with fewer points
    than subsamples some tag is missing)."""
    n = _n(t)
    if N_samples > n:
        return False
    if isinstance(t, torch.Tensor):
        return bool(torch.bincount(t.reshape(-1), minlength=N_samples + 1)[1:].min() > 0)
    return bool(np.bincount(np.ravel(t), minlength=N_samples + 1)[1:].min() > 0)


def _process_weights_jtags(sample1, sample2, weights1, weights2, jtags1, jtags2, N_samples):
    """PC/npairs_jackknife_3d.py:217-276 (identical in PC/npairs_jackknife_xy_z.py:202-261)."""
    weights1 = _as_weights(weights1, sample1, "weights1")
    weights2 = _as_weights(weights2, sample2, "weights2")
    jtags1 = _as_tags(jtags1, sample1, "jtags1")
    jtags2 = _as_tags(jtags2, sample2, "jtags2")
    if _min(jtags1) < 1:
        raise HalotoolsError("jtags1 must be >= 1")
    if _min(jtags2) < 1:
        raise HalotoolsError("jtags2 must be >= 1")
    if _max(jtags1) > N_samples:
        raise HalotoolsError("jtags1 must be <= N_samples")
    if _max(jtags2) > N_samples:
        raise HalotoolsError("jtags2 must be <= N_samples")
    covers = _covers_all(jtags1, N_samples)
    if not covers:
        warn("Warning: sample1 does not contain points in every jackknife sample.")
    # the reference tests jtags1 again for its sample2 warning (PC/npairs_jackknife_3d.py:273-274)
    if not covers:
        warn("Warning: sample2 does not contain points in every jackknife sample.")
    return weights1, weights2, jtags1, jtags2


# the reference's names for the weights / tags processing
_npairs_jackknife_3d_process_weights_jtags = _process_weights_jtags
_npairs_jackknife_xy_z_process_weights_jtags = _process_weights_jtags


def _same_values(a, b):
    """Known to hold the same values: one object, one view of a device tensor
    (not compared by value, which would wait for the device), or equal host
    arrays.  A False only costs the symmetric evaluation of an auto-count."""
    if a is b:
        return True
    if a is None or b is None:
        return False
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        return D.same_array(a, b)
    return np.array_equal(a, b)


def _run(kind, sample1, sample2, cols1, cols2, period, pbc, r2, p2, weights1, weights2, jtags1, jtags2, N_samples,
         raw_same=False):
    if N_samples >= 2 ** 31:
        raise ValueError("N_samples must be below 2**31")
    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    # raw_same: the caller's weights and tags arguments were already known equal
    same = D.same_array(sample1, sample2) and (raw_same or (_same_values(weights1, weights2)
                                                            and _same_values(jtags1, jtags2)))
    s1, s2 = C.upload(cols1, cols2, D.same_array(sample1, sample2), device)
    tag = lambda t: (t if isinstance(t, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(t))).to(
        device=device, dtype=torch.int32).contiguous()
    jt1 = tag(jtags1)
    jt2 = jt1 if same else tag(jtags2)
    w1 = D.to_device_weights(weights1, device).contiguous() if weights1 is not None else None
    w2 = None if same else (D.to_device_weights(weights2, device).contiguous() if weights2 is not None else None)
    mock = C.mock_dict(len(s1[0]), len(s2[0]), period, pbc)
    res = D.run(kind, [mock], s1, s2, r2, p2=p2, same=same, w1=w1, w2=w2,
                nw=1 if (w1 is not None or w2 is not None) else 0, jtags1=jt1, jtags2=jt2,
                n_samples=N_samples, device=device)
    status = int(res.cbuf[-1].item()) & 0xFFFFFFFF
    if status & 2:
        raise HalotoolsError("jackknife tags must lie in [1, N_samples]")
    if status:
        if pbc:
            raise ValueError("All points must lie inside the periodic box [0, period] in every dimension "
                             "and be finite")
        raise ValueError("All point coordinates must be finite")
    return res.sums[0] if as_tensor else C.to_host(res.sums[0])


def npairs_jackknife_3d(sample1, sample2, rbins, jtags1, jtags2, N_samples, period=None, weights1=None,
                        weights2=None, num_threads=1, approx_cell1_size=None, approx_cell2_size=None):
    """Jackknife pair sums, float64 of shape (N_samples + 1, len(rbins)) (NumPy, or
    a CUDA tensor when ``sample1`` is one), as the reference (PC/npairs_jackknife_3d.py:24-214).

    Row 0 sums w1*w2 over all pairs within each edge; row s the same with
    subsample s removed (pairs with one point in s count half).
    """
    # PC/npairs_jackknife_3d.py:154-214: the argument processing of npairs_3d first
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
    raw_same = _same_values(weights1, weights2) and _same_values(jtags1, jtags2)
    weights1, weights2, jtags1, jtags2 = _process_weights_jtags(sample1, sample2, weights1, weights2, jtags1,
                                                                jtags2, N_samples)
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rmax, rmax, rmax), (xperiod, yperiod, zperiod))
    return _run(E.HP_NPAIRS_JACKKNIFE_3D, sample1, sample2, (x1, y1, z1), (x2, y2, z2), period, pbc,
                P.squared(rbins), None, weights1, weights2, jtags1, jtags2, N_samples, raw_same)


def npairs_jackknife_xy_z(sample1, sample2, rp_bins, pi_bins, jtags1, jtags2, N_samples, period=None,
                          weights1=None, weights2=None, num_threads=1, approx_cell1_size=None,
                          approx_cell2_size=None):
    """Jackknife (rp, pi) pair sums, float64 of shape (N_samples + 1, len(rp_bins),
    len(pi_bins)), as the reference (PC/npairs_jackknife_xy_z.py:23-199)."""
    # PC/npairs_jackknife_xy_z.py:153-199: the argument processing of npairs_xy_z first
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
    raw_same = _same_values(weights1, weights2) and _same_values(jtags1, jtags2)
    weights1, weights2, jtags1, jtags2 = _process_weights_jtags(sample1, sample2, weights1, weights2, jtags1,
                                                                jtags2, N_samples)
    P.set_approximate_cell_sizes(approx_cell1_size, approx_cell2_size, period)
    P.check_search_lengths((rp_max, rp_max, pi_max), (xperiod, yperiod, zperiod))
    return _run(E.HP_NPAIRS_JACKKNIFE_XY_Z, sample1, sample2, (x1, y1, z1), (x2, y2, z2), period, pbc,
                P.squared(rp_bins), P.squared(pi_bins), weights1, weights2, jtags1, jtags2, N_samples, raw_same)
