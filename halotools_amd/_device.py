"""Device-side plumbing shared by the drop-in pair counters.

PyTorch-ROCm is used only for device memory and the current HIP stream; the
counting itself is libhalopair.so (hand-written HIP kernels behind a C ABI).
"""
import ctypes

import numpy as np
import torch

from . import _engine as E


def require_gpu(device=None):
    if not torch.cuda.is_available():
        raise RuntimeError("halotools_amd pair counters need a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    E.library()
    return torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())


def is_device_tensor(a):
    return isinstance(a, torch.Tensor) and a.is_cuda


def same_array(a, b):
    """True when a and b are the same sample (identical object or identical view)."""
    if a is b:
        return True
    if isinstance(a, np.ndarray) and isinstance(b, np.ndarray):
        return (a.shape == b.shape and a.strides == b.strides and a.dtype == b.dtype
                and a.__array_interface__["data"][0] == b.__array_interface__["data"][0])
    if isinstance(a, torch.Tensor) and isinstance(b, torch.Tensor):
        return (a.shape == b.shape and a.stride() == b.stride() and a.dtype == b.dtype
                and a.device == b.device and a.data_ptr() == b.data_ptr())
    return False


def xyz_rows(cols):
    """The C-ordered float64 (N, 3) array whose columns ``cols`` are, or None.

    ``sample[:, 0], sample[:, 1], sample[:, 2]`` of a C-contiguous float64
    sample are column views 8 bytes apart with a 24-byte stride; the rows
    they come from are then one contiguous block that can be copied as is.
    """
    if not all(isinstance(c, np.ndarray) for c in cols):
        return None
    c0 = cols[0]
    n = len(c0)
    if n == 0 or any(c.dtype != np.float64 or c.ndim != 1 or len(c) != n for c in cols):
        return None
    if any(c.strides != (24,) for c in cols):
        return None
    p0 = c0.__array_interface__["data"][0]
    if any(cols[k].__array_interface__["data"][0] != p0 + 8 * k for k in range(3)):
        return None
    if not all(c.flags.writeable for c in cols):
        return None
    return np.lib.stride_tricks.as_strided(c0, shape=(n, 3), strides=(24, 8))


def tensor_xyz_rows(cols):
    """The float64 (N, 3) tensor whose columns ``cols`` are (as a view), or None."""
    if len(cols) != 3:
        return None
    c0, c1, c2 = cols
    if not (isinstance(c0, torch.Tensor) and isinstance(c1, torch.Tensor) and isinstance(c2, torch.Tensor)):
        return None
    f64 = torch.float64
    if c0.dtype != f64 or c1.dtype != f64 or c2.dtype != f64:
        return None
    # a stride of (3,) also means one dimension
    if c0.stride() != (3,) or c1.stride() != (3,) or c2.stride() != (3,):
        return None
    n = c0.shape[0]
    if n == 0 or c1.shape[0] != n or c2.shape[0] != n or c1.device != c0.device or c2.device != c0.device:
        return None
    p0 = c0.data_ptr()
    if c1.data_ptr() != p0 + 8 or c2.data_ptr() != p0 + 16:
        return None
    return torch.as_strided(c0, (n, 3), (3, 1))


class RowColumns(tuple):
    """The x, y, z column views of one (N, 3) row-major float64 tensor."""
    stride = 3


def _row_columns(rows):
    """The x, y, z column views of (N, 3) float64 device rows (the library reads
    them in place with a stride of 3 doubles)."""
    return RowColumns(rows.unbind(1))


def to_device_columns(cols, device):
    """Three coordinate columns (NumPy or torch, any real dtype) -> float64 device
    tensors: contiguous columns, or the three column views of one (N, 3) array."""
    if all(isinstance(c, torch.Tensor) for c in cols):
        rows = tensor_xyz_rows(cols)
        if rows is not None:
            if rows.device == device:
                return RowColumns(cols)  # already the column views of rows on the device
            return _row_columns(rows.to(device))
        return tuple(c.to(device=device, dtype=torch.float64).contiguous() for c in cols)
    base = xyz_rows(cols)
    if base is not None:
        # one copy of the caller's (N, 3) rows, read in place on the device
        return _row_columns(torch.from_numpy(base).to(device, non_blocking=False))
    host = np.empty((3, len(cols[0])), dtype=np.float64)
    for k in range(3):
        host[k] = np.asarray(cols[k])
    t = torch.from_numpy(host).to(device, non_blocking=False)
    return t[0], t[1], t[2]


def to_device_weights(w, device):
    if isinstance(w, torch.Tensor):
        return w.to(device=device, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(w, dtype=np.float64)).to(device)


def _coordinate_columns(cols):
    """(columns, stride) for a request: contiguous columns (stride 1), the
    column views of one (N, 3) row-major tensor (stride 3), or contiguous
    copies of anything else."""
    if isinstance(cols, RowColumns):
        return cols, 3
    if all(c.is_contiguous() for c in cols):
        return cols, 1
    if tensor_xyz_rows(cols) is not None:
        return cols, 3
    return tuple(c.contiguous() for c in cols), 1


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Result(object):
    pass


def run(*args, **kwargs):
    """Enqueue one request on the current stream and return the output tensors (see ``prepare``)."""
    res = prepare(*args, **kwargs)
    launch(res)
    return res


def launch(res, stream=None):
    """(Re)issue a prepared request on `stream` (default: the current stream of its device)."""
    req = res.req
    req.stream = stream if stream is not None else torch.cuda.current_stream(res.device).cuda_stream
    E.check(res.fn(ctypes.byref(req)))


def prepare(kind, mocks, s1, s2, r2, p2=None, same=False, w1=None, w2=None, nw=0, wfunc=0,
            want_counts=True, want_sums=False, rows1=False, rows2=False, want_evals=False,
            also_counts=False, device=None, pair_events=None, build_events=None, tile_waves=0, tangent=False,
            jtags1=None, jtags2=None, n_samples=0):
    """Build a request with all of its output and scratch buffers allocated.

    mocks : list of dicts with n1, n2, period (3,), pbc, cell (3,) (cell <= 0: default)
    s1, s2 : tuples (x, y, z) of float64 device tensors, mocks concatenated
    r2 : host float64 array of squared radial / rp edges (already squared like the reference)
    pair_events, build_events : optional (start, end) HIP event handles recorded
        around the pair kernel(s) and around the cell build
    tile_waves : wavefronts per pair-kernel workgroup (1, 2, 4, 8; 0 = library default)
    tangent : marked id 1 with (w, dw/dtheta) weights: also res.tangents, d(sums)/dtheta
    jtags1, jtags2, n_samples : jackknife kinds: int32 device tags (1..n_samples);
        w1 / w2 are then optional (n,) weights and res.sums has shape
        (M, n_samples + 1, nr) or (M, n_samples + 1, nr, np)
    The returned object keeps every buffer the request points to alive; the
    same request can be launched again (e.g. inside HIP-graph capture) after
    its input buffers are refreshed in place.
    """
    device = torch.device(device) if device is not None else s1[0].device
    M = len(mocks)
    marr = (E.HpMock * M)()
    n1_tot = n2_tot = 0
    for m, d in enumerate(mocks):
        marr[m].n1 = int(d["n1"])
        marr[m].n2 = int(d["n1"] if same else d["n2"])
        marr[m].period[:] = [float(v) for v in d["period"]]
        marr[m].pbc = 1 if d["pbc"] else 0
        marr[m].cell[:] = [float(v) for v in d.get("cell", (0.0, 0.0, 0.0))]
        n1_tot += marr[m].n1
        n2_tot += marr[m].n2
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    nr = len(r2)
    if p2 is not None:
        p2 = np.ascontiguousarray(p2, dtype=np.float64)
    flags = 0
    if same:
        flags |= E.HP_SAME_SAMPLES
    marked = kind in (E.HP_MARKED_NPAIRS_3D, E.HP_MARKED_NPAIRS_XY_Z)
    if marked:
        if rows1:
            flags |= E.HP_ROWS1
        if rows2:
            flags |= E.HP_ROWS2
        if not want_sums:
            flags |= E.HP_NO_SUMS
        if also_counts:
            flags |= E.HP_ALSO_COUNTS
        if tangent:
            flags |= E.HP_TANGENT
    req = E.HpRequest()
    req.kind = kind
    req.flags = flags
    req.nmocks = M
    req.wfunc = int(wfunc)
    req.mocks = ctypes.cast(marr, ctypes.POINTER(E.HpMock))
    s1, req.stride1 = _coordinate_columns(s1)
    req.x1, req.y1, req.z1 = [c.data_ptr() for c in s1]
    if not same:
        s2, req.stride2 = _coordinate_columns(s2)
        req.x2, req.y2, req.z2 = [c.data_ptr() for c in s2]
    req.w1 = w1.data_ptr() if w1 is not None else None
    req.w2 = w2.data_ptr() if (w2 is not None and not same) else None
    req.nw = int(nw)
    req.nr = nr
    req.np = len(p2) if p2 is not None else 0
    req.r2 = r2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    req.p2 = p2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if p2 is not None else None

    res = Result()
    obj = kind == E.HP_NPAIRS_PER_OBJECT_3D
    ncount = nr * (len(p2) if kind == E.HP_NPAIRS_XY_Z else 1)
    need_counts = (not marked and kind not in (E.HP_NPAIRS_JACKKNIFE_3D, E.HP_NPAIRS_JACKKNIFE_XY_Z)) or also_counts
    # counts and the status word share one buffer so one copy brings both back
    clen = n1_tot * nr if obj else M * ncount
    cbuf = torch.empty(clen + 1 if need_counts else 1, dtype=torch.int64, device=device)
    res.status = cbuf[-1:]
    if need_counts:
        shape = (M, nr, len(p2)) if kind == E.HP_NPAIRS_XY_Z else ((n1_tot, nr) if obj else (M, nr))
        res.counts = cbuf[:-1].view(*shape)
        req.counts = cbuf.data_ptr()
    else:
        res.counts = None
    req.status = res.status.data_ptr()
    jack = kind in (E.HP_NPAIRS_JACKKNIFE_3D, E.HP_NPAIRS_JACKKNIFE_XY_Z)
    sum_shape = (M, nr, len(p2)) if kind == E.HP_MARKED_NPAIRS_XY_Z else (M, nr)
    if jack:
        sum_shape = (M, int(n_samples) + 1, nr) + ((len(p2),) if kind == E.HP_NPAIRS_JACKKNIFE_XY_Z else ())
        req.jtags1 = jtags1.data_ptr()
        req.jtags2 = jtags2.data_ptr() if (jtags2 is not None and not same) else None
        req.n_samples = int(n_samples)
    res.sums = (torch.empty(sum_shape, dtype=torch.float64, device=device)
                if ((marked and want_sums) or jack) else None)
    req.sums = res.sums.data_ptr() if res.sums is not None else None
    res.tangents = torch.empty((M, nr), dtype=torch.float64, device=device) if (marked and tangent) else None
    req.tangents = res.tangents.data_ptr() if res.tangents is not None else None
    res.rows1 = torch.empty((n1_tot, nr), dtype=torch.float64, device=device) if rows1 else None
    res.rows2 = torch.empty((n2_tot, nr), dtype=torch.float64, device=device) if rows2 else None
    req.rows1 = res.rows1.data_ptr() if rows1 else None
    req.rows2 = res.rows2.data_ptr() if rows2 else None
    res.evals = torch.empty(M, dtype=torch.int64, device=device) if want_evals else None
    req.evals = res.evals.data_ptr() if want_evals else None

    lib = E.library()
    nbytes = lib.hp_workspace_size(ctypes.byref(req))
    if nbytes == 0:
        E.check(-1)
    ws = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
    req.workspace = ws.data_ptr()
    req.workspace_bytes = int(nbytes)
    if pair_events is not None:
        req.pair_events[0], req.pair_events[1] = pair_events
    if build_events is not None:
        req.build_events[0], req.build_events[1] = build_events
    req.tile_waves = int(tile_waves)
    res.fn = {E.HP_NPAIRS_3D: lib.hp_npairs_3d, E.HP_NPAIRS_PROJECTED: lib.hp_npairs_projected,
              E.HP_NPAIRS_XY_Z: lib.hp_npairs_xy_z, E.HP_MARKED_NPAIRS_3D: lib.hp_marked_npairs_3d,
              E.HP_NPAIRS_PER_OBJECT_3D: lib.hp_npairs_per_object_3d,
              E.HP_MARKED_NPAIRS_XY_Z: lib.hp_marked_npairs_xy_z,
              E.HP_NPAIRS_JACKKNIFE_3D: lib.hp_npairs_jackknife_3d,
              E.HP_NPAIRS_JACKKNIFE_XY_Z: lib.hp_npairs_jackknife_xy_z}[kind]
    res.req = req
    res.device = device
    # keep every host and device buffer the request points to alive with it
    res.keep = (marr, r2, p2, s1, s2, w1, w2, jtags1, jtags2)
    res.workspace = ws
    res.cbuf = cbuf
    return res


def backward_rows(rows, gsuffix, offsets, out=None):
    """grad[p] = rows[p] . gsuffix[mock(p)] on the current stream."""
    n = int(offsets[-1])
    grad = out if out is not None else torch.empty(n, dtype=torch.float64, device=rows.device)
    offs = (ctypes.c_int64 * len(offsets))(*[int(v) for v in offsets])
    g = gsuffix.contiguous()
    E.check(E.library().hp_marked_backward(
        ctypes.c_void_p(rows.data_ptr()), rows.shape[1], ctypes.c_void_p(g.data_ptr()), offs,
        len(offsets) - 1, ctypes.c_void_p(grad.data_ptr()),
        ctypes.c_void_p(torch.cuda.current_stream(rows.device).cuda_stream)))
    return grad


def zheng07_tangent(log_mass, mock_id, logMmin, sigma, out):
    """out[:, 0] = Zheng07 central weights, out[:, 1] = their logMmin
    derivatives, per point of mock mock_id[i] (one library kernel; the same
    values as the elementwise torch expression)."""
    n = log_mass.numel()
    if out.shape != (n, 2) or not out.is_contiguous() or out.dtype != torch.float64:
        raise ValueError("out must be a contiguous float64 (n, 2) tensor")
    if log_mass.dtype != torch.float64 or mock_id.dtype != torch.int64 or logMmin.dtype != torch.float64:
        raise ValueError("log_mass and logMmin must be float64, mock_id int64")
    if not (log_mass.is_contiguous() and mock_id.is_contiguous() and logMmin.is_contiguous()):
        raise ValueError("inputs must be contiguous")
    if mock_id.numel() != n:
        raise ValueError("mock_id must have one entry per point")
    E.check(E.library().hp_zheng07_tangent(
        ctypes.c_void_p(log_mass.data_ptr()), ctypes.c_void_p(mock_id.data_ptr()),
        ctypes.c_void_p(logMmin.data_ptr()), float(sigma), n, ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(torch.cuda.current_stream(out.device).cuda_stream)))
    return out


class SegmentSum(object):
    """Per-mock sums out[m] = values[offsets[m]:offsets[m+1]].sum() in a fixed
    order (library kernels, no atomics), with the scratch allocated once so the
    call can be captured in a HIP graph."""

    def __init__(self, offsets, device):
        self.offsets = (ctypes.c_int64 * len(offsets))(*[int(v) for v in offsets])
        self.nmocks = len(offsets) - 1
        lib = E.library()
        need = lib.hp_segment_sum_scratch(self.offsets, self.nmocks)
        if need < 0:
            E.check(-1)
        self.scratch = torch.empty(int(need), dtype=torch.float64, device=device)

    def __call__(self, values, out):
        E.check(E.library().hp_segment_sum(
            ctypes.c_void_p(values.data_ptr()), self.offsets, self.nmocks, ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(self.scratch.data_ptr()), self.scratch.numel(),
            ctypes.c_void_p(torch.cuda.current_stream(values.device).cuda_stream)))
        return out
