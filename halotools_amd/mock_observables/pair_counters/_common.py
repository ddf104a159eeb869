"""Shared flow of the drop-in pair counters: samples to the device, one engine
call, results back in the caller's container (NumPy in -> NumPy out; CUDA
tensor in -> CUDA tensor out)."""
import numpy as np
import torch

from ... import _device as D
from . import _process as P


def prepare_samples(sample1, sample2, period, min_size):
    """Column views, the non-PBC box enclosure and the periodic box, then upload.

    Returns (s1, s2, same, period3, pbc, device, as_tensor).
    """
    device = D.require_gpu(sample1.device if D.is_device_tensor(sample1) else None)
    as_tensor = D.is_device_tensor(sample1)
    same = D.same_array(sample1, sample2)
    x1, y1, z1 = P.columns(sample1)
    x2, y2, z2 = P.columns(sample2)
    if period is None:
        pbc = False
        x1, y1, z1, x2, y2, z2, period = P.enclose_in_box(x1, y1, z1, x2, y2, z2, min_size=min_size)
    else:
        pbc = True
        period = P.process_period(period)
    return (x1, y1, z1), (x2, y2, z2), same, period, pbc, device, as_tensor


def upload(cols1, cols2, same, device):
    s1 = D.to_device_columns(cols1, device)
    s2 = s1 if same else D.to_device_columns(cols2, device)
    return s1, s2


def mock_dict(n1, n2, period, pbc, cell=None):
    return dict(n1=int(n1), n2=int(n2), period=[float(v) for v in period[:3]], pbc=bool(pbc),
                cell=list(cell) if cell is not None else [0.0, 0.0, 0.0])


def finish(res, out, as_tensor, period, pbc):
    """Check the status word and hand the result back."""
    if as_tensor:
        status = int(res.cbuf[-1].item()) & 0xFFFFFFFF
        host = None
    else:
        host = res.cbuf.cpu().numpy()
        status = int(host[-1]) & 0xFFFFFFFF
    if status:
        if pbc:
            raise ValueError("All points must lie inside the periodic box [0, period] in every dimension "
                             "and be finite")
        raise ValueError("All point coordinates must be finite")
    return out(host)


def to_host(t):
    return t.detach().cpu().numpy()
