"""Cell-size autotuner for the pair kernels.

The GPU mesh has one free parameter per realisation: the edge of the fine
stream-side cells.  Finer cells tighten each group's window (fewer candidate
pairs), coarser cells cut the per-segment overhead.  The best edge depends on
the point density, the search radius and the kernel mode, so the tuner times
the count kernel on the caller's own points for a few targets of "stream points
per cell" and keeps the fastest, keyed by (mode, log2 density bucket, search
radius / box bucket).  Results are cached in memory and, when
HALOTOOLS_AMD_TUNE_CACHE names a file, in that JSON file.
"""
import json
import math
import os
import time

import numpy as np
import torch

from . import _device as D
from . import _engine as E

TARGETS = (2.0, 4.0, 8.0, 16.0, 32.0)
_CACHE = {}


def _cache_path():
    return os.environ.get("HALOTOOLS_AMD_TUNE_CACHE")


def _load():
    p = _cache_path()
    if p and os.path.exists(p) and not _CACHE:
        try:
            _CACHE.update(json.load(open(p)))
        except (OSError, ValueError):
            pass


def _save():
    p = _cache_path()
    if p:
        with open(p, "w") as f:
            json.dump(_CACHE, f, indent=1, sort_keys=True)


def key(kind, n, period, rmax):
    vol = float(np.prod(period))
    dens = max(n, 1) / vol
    return "%d:%d:%d" % (kind, int(round(math.log2(dens))), int(round(4 * math.log2(rmax / min(period)))))


def cell_edge(target, n, period, rmax):
    """Cell edge giving about `target` stream points per cell, clamped to [rmax/3, rmax]."""
    dens = max(n, 1) / float(np.prod(period))
    return float(min(max((target / dens) ** (1.0 / 3.0), rmax / 3.0), rmax))


def lookup(kind, n, period, rmax):
    """Tuned cell edge for this shape of problem, or None (library default)."""
    _load()
    t = _CACHE.get(key(kind, n, period, rmax))
    return None if t is None else cell_edge(t, n, period, rmax)


def tune(points, period, rbins, reps=2, targets=TARGETS, device=None):
    """Time npairs_3d auto-counts of `points` for each target and cache the fastest.

    points : (N, 3) array or CUDA tensor inside the periodic box `period`.
    Returns (best_target, {target: seconds}).
    """
    device = D.require_gpu(device)
    t = points if isinstance(points, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(points))
    t = t.to(device=device, dtype=torch.float64)
    cols = (t[:, 0].contiguous(), t[:, 1].contiguous(), t[:, 2].contiguous())
    period = np.atleast_1d(period).astype(float)
    if len(period) == 1:
        period = np.array([period[0]] * 3)
    rbins = np.asarray(rbins, dtype=np.float64)
    rmax = float(np.max(rbins))
    n = len(t)
    timing = {}
    ref = None
    for target in targets:
        edge = cell_edge(target, n, period, rmax)
        mock = dict(n1=n, n2=n, period=list(period), pbc=True, cell=[edge] * 3)
        best = float("inf")
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = D.run(E.HP_NPAIRS_3D, [mock], cols, cols, rbins * rbins, same=True, device=device)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        counts = res.counts[0].cpu().numpy()
        if ref is None:
            ref = counts
        elif not np.array_equal(ref, counts):
            raise RuntimeError("counts changed with the cell size; this is a bug")
        timing[target] = best
    winner = min(timing, key=timing.get)
    _load()
    _CACHE[key(E.HP_NPAIRS_3D, n, period, rmax)] = winner
    _save()
    return winner, timing
