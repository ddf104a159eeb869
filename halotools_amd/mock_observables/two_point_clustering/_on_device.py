"""Device residency of the two-point caller layer.

The reference's two-point functions call the pair counters once per needed
combination (DD, DR, RR), difference the cumulative counts with ``np.diff``
and run an estimator, all in NumPy.  Here the samples go to the GPU once,
every pair counter runs on the device tensors (its result stays in HBM), the
differences and the estimator run on the device, and only the final
correlation function comes back -- to NumPy when the caller passed NumPy
arrays, as a CUDA tensor when the caller passed CUDA tensors.
"""
import numpy as np
import torch

from ... import _device as D


class Residency(object):
    """Where the inputs live and how results go back."""

    def __init__(self, sample1):
        self.host_out = not D.is_device_tensor(sample1)
        self.device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)

    def put(self, a):
        """A sample / marks array on the device, with its dtype and values unchanged."""
        if a is None:
            return None
        if isinstance(a, torch.Tensor):
            return a.to(self.device)
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)

    def put_f64(self, a):
        if a is None:
            return None
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=torch.float64)
        return torch.as_tensor(np.asarray(a, dtype=np.float64), device=self.device)

    def out(self, x):
        if x is None or not self.host_out:
            return x
        h = x.detach().cpu().numpy()
        return h[()] if h.ndim == 0 else h


def diff1(c):
    """np.diff of cumulative counts (int64 or float64) on the device, as float64."""
    return torch.diff(c).to(torch.float64)


def diff2(c):
    """np.diff(np.diff(c, axis=0), axis=1) on the device, as float64."""
    return torch.diff(torch.diff(c, dim=0), dim=1).to(torch.float64)
