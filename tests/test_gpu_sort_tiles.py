"""GPU cell build with both counting-sort tile sizes.

The build sorts cell keys with a stable LSD counting sort in tiles of 2048
keys, or 8192 keys for large sorts (the reference's mesh is a NumPy argsort,
PC/rectangular_mesh.py:211-222, PC = halotools/mock_observables/pair_counters).
A stable sort has one result whatever the tile size, so forcing the large
tiles on small inputs (HALOPAIR_SORT_TILE, read at every build) must give the
same sorted order, hence bitwise the same counts, per-point rows and marked
sums as the default tiles; the counts are also checked against the oracle.
Sizes straddle the tile sizes (ragged last tiles, a single partial tile).
"""
import os

import numpy as np
import pytest
import torch

from oracle import mesh as ORC
from tests.helpers import random_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mo(gpu):
    import halotools_amd.mock_observables as mo
    return mo


def _with_tile(tile, fn):
    os.environ["HALOPAIR_SORT_TILE"] = str(tile)
    try:
        return fn()
    finally:
        del os.environ["HALOPAIR_SORT_TILE"]


@pytest.mark.parametrize("n1,n2", [(5000, None), (8191, 8193), (16385, None), (20000, 3000)])
def test_counts_and_rows_tile_invariant(mo, n1, n2):
    s1 = random_sample(n1, seed=n1)
    s2 = s1 if n2 is None else random_sample(n2, seed=n2)
    rb = np.logspace(-3, np.log10(0.08), 12)
    w1 = np.random.default_rng(n1).uniform(0.5, 1.5, len(s1))
    w2 = w1 if n2 is None else np.random.default_rng(n2).uniform(0.5, 1.5, len(s2))

    def run():
        c = mo.npairs_3d(s1, s2, rb, period=1.0)
        o = mo.npairs_per_object_3d(s1, s2, rb, period=1.0)
        w = mo.marked_npairs_3d(s1, s2, rb, period=1.0, weights1=w1, weights2=w2, weight_func_id=1)
        return c, o, w

    small = _with_tile(2048, run)
    large = _with_tile(8192, run)
    for a, b in zip(small, large):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert np.array_equal(small[0], ORC.npairs_3d(s1, s2, rb, period=1.0, num_threads=8))


def test_batch_tile_invariant(gpu):
    """Uneven mocks in one sort (per-mock key offsets) with both tile sizes."""
    from halotools_amd import batch
    rng = np.random.default_rng(7)
    sizes = [9000, 700, 12000]
    L = 100.0
    pts = [torch.tensor(rng.uniform(0, L, (n, 3)), device=gpu) for n in sizes]
    lm = torch.tensor(rng.uniform(11, 15, sum(sizes)), device=gpu)
    rb = np.logspace(-1, np.log10(12.0), 15)
    c = torch.tensor(rng.normal(size=15), device=gpu)
    logMmin = torch.tensor([11.8, 12.1, 12.3], device=gpu)

    def run():
        b = batch.MockBatch(pts, L, rb, device=gpu)
        return [x.clone() for x in batch.HODBatchStep(b, lm, logMmin, 0.3, c).run()]

    small = _with_tile(2048, run)
    large = _with_tile(8192, run)
    for a, b in zip(small, large):
        assert torch.equal(a, b)
    for m, p in enumerate(pts):
        ph = p.cpu().numpy()
        assert np.array_equal(small[0][m].cpu().numpy(), ORC.npairs_3d(ph, ph, rb, period=L, num_threads=8))


@pytest.mark.parametrize("tile", [2048, 8192])
def test_batch_with_empty_mocks(gpu, tile):
    """Mocks without points between others: their tile ranges are empty, and
    every other mock's counts still equal the oracle's (multi-mock requests
    through the C ABI, auto and cross)."""
    from halotools_amd import _device as D, _engine as E
    rng = np.random.default_rng(tile)
    L = 100.0
    rb = np.logspace(-1, np.log10(12.0), 9)
    sizes1 = [0, 9000, 0, 700, 12000, 0]
    sizes2 = [500, 8000, 0, 0, 3000, 0]
    p1 = [rng.uniform(0, L, (n, 3)) for n in sizes1]
    p2 = [rng.uniform(0, L, (n, 3)) for n in sizes2]

    def cols(ps):
        t = torch.tensor(np.concatenate(ps), device=gpu)
        return tuple(t[:, i].contiguous() for i in range(3))

    c1, c2 = cols(p1), cols(p2)

    def run():
        auto = D.run(E.HP_NPAIRS_3D, [dict(n1=a, n2=a, period=[L] * 3, pbc=True) for a in sizes1], c1, c1, rb * rb,
                     same=True, device=gpu)
        cross = D.run(E.HP_NPAIRS_3D, [dict(n1=a, n2=b, period=[L] * 3, pbc=True) for a, b in zip(sizes1, sizes2)],
                      c1, c2, rb * rb, device=gpu)
        return auto.counts.cpu().numpy(), cross.counts.cpu().numpy()

    auto, cross = _with_tile(tile, run)
    for m in range(len(sizes1)):
        want = ORC.npairs_3d(p1[m], p1[m], rb, period=L, num_threads=8) if sizes1[m] else np.zeros(len(rb), np.int64)
        assert np.array_equal(auto[m], want), m
        both = sizes1[m] and sizes2[m]
        want = ORC.npairs_3d(p1[m], p2[m], rb, period=L, num_threads=8) if both else np.zeros(len(rb), np.int64)
        assert np.array_equal(cross[m], want), m
