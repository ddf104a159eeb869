"""Autotuner for the pair kernels: fine-cell size and workgroup tile shape.

The GPU mesh has one free parameter per realisation, the edge of the fine
stream-side cells: finer cells tighten each lane group's window (fewer
candidate pairs), coarser cells cut the per-segment overhead.  The pair kernel
has one more, the tile shape: how many 64-point lane groups (wavefronts) share
a workgroup, which trades LDS per workgroup against waves in flight per CU.
The best choice depends on the point density, the search radius and the kernel
mode, so the tuner times the caller's own points for a grid of "stream points
per cell" targets and workgroup sizes and keeps the fastest, keyed by (mode,
log2 density bucket, search radius / box bucket).  Results are cached in
memory and, when HALOTOOLS_AMD_TUNE_CACHE names a file, in that JSON file.

Modes: ``E.HP_NPAIRS_3D`` times npairs_3d auto-counts;
``E.HP_MARKED_NPAIRS_3D`` times the product-weighted pass that produces both
samples' Jacobian rows (reverse mode); ``TANGENT`` times the forward-mode pass
with counts (the batched model-fitting step).
"""
import json
import math
import os
import time

import numpy as np
import torch

from . import _device as D
from . import _engine as E

TARGETS = (8.0, 32.0, 128.0, 4096.0)
TANGENT = 1000  # tuner key of the marked tangent pass (not a library kind)
WAVES = (2, 4, 8)
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
    """Tuned (cell edge, workgroup waves) for this shape of problem, or None (library defaults)."""
    _load()
    t = _CACHE.get(key(kind, n, period, rmax))
    if t is None:
        return None
    return cell_edge(float(t["target"]), n, period, rmax), int(t["waves"])


def _request(kind, n, period, cols, rbins, edge, waves, device):
    mock = dict(n1=n, n2=n, period=list(period), pbc=True, cell=[edge] * 3)
    r2 = rbins * rbins
    if kind == E.HP_NPAIRS_3D:
        return D.prepare(kind, [mock], cols, cols, r2, same=True, device=device, tile_waves=waves)
    if kind == TANGENT:
        wt = torch.ones((n, 2), dtype=torch.float64, device=device)
        return D.prepare(E.HP_MARKED_NPAIRS_3D, [mock], cols, cols, r2, same=True, w1=wt, nw=2, wfunc=1,
                         want_counts=False, want_sums=True, also_counts=True, tangent=True, device=device,
                         tile_waves=waves)
    w = torch.ones((n, 1), dtype=torch.float64, device=device)
    return D.prepare(kind, [mock], cols, cols, r2, same=True, w1=w, nw=1, wfunc=1, want_counts=False,
                     want_sums=True, rows1=True, rows2=True, device=device, tile_waves=waves)


def tune(points, period, rbins, kind=E.HP_NPAIRS_3D, reps=2, targets=TARGETS, waves=WAVES, device=None):
    """Time auto-counts of `points` for every (target, waves) and cache the fastest.

    points : (N, 3) array or CUDA tensor inside the periodic box `period`.
    kind   : E.HP_NPAIRS_3D, E.HP_MARKED_NPAIRS_3D or TANGENT (see the module docstring).
    Returns ((best_target, best_waves), {(target, waves): seconds}).  The counts
    of every configuration are checked equal: the cell size and the tile shape
    change the work schedule, never the result.
    """
    if kind not in (E.HP_NPAIRS_3D, E.HP_MARKED_NPAIRS_3D, TANGENT):
        raise ValueError("tune: kind must be HP_NPAIRS_3D, HP_MARKED_NPAIRS_3D or TANGENT")
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
        for wv in waves:
            res = _request(kind, n, period, cols, rbins, edge, wv, device)
            best = float("inf")
            for _ in range(reps + 1):
                torch.cuda.synchronize(device)
                t0 = time.perf_counter()
                D.launch(res)
                torch.cuda.synchronize(device)
                best = min(best, time.perf_counter() - t0)
            counts = res.counts[0].cpu().numpy() if res.counts is not None else res.sums[0].cpu().numpy()
            if ref is None:
                ref = counts
            elif not np.array_equal(ref, counts):
                if res.counts is not None:
                    raise RuntimeError("counts changed with the cell size or tile shape; this is a bug")
                if not np.allclose(ref, counts, rtol=1e-12, atol=0):
                    raise RuntimeError("sums changed with the cell size or tile shape; this is a bug")
            timing[(target, wv)] = best
            del res
    winner = min(timing, key=timing.get)
    _load()
    _CACHE[key(kind, n, period, rmax)] = {"target": winner[0], "waves": winner[1]}
    _save()
    return winner, timing
