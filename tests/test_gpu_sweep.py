"""GPU parity over a seeded sweep of shapes the fixed cases do not pin one by one.

Each case draws its own sample sizes (from empty-ish to a few thousand
points), box (cubic or not, periodic or enclosed), bin edges (1 to 80 edges,
so both the per-lane and the shared per-wave count histograms run), cell-size
hints and point layout (uniform, or clumped so that lane groups see dense
and empty cells), and runs npairs_3d, npairs_projected, npairs_xy_z,
npairs_per_object_3d, the product-marked sums (3-D and (rp, pi)) and the
jackknife tables (3-D and (rp, pi), random subsample tags) through the
drop-in API against the oracle: counts and unit-weight jackknife tables
bit-exact, marked sums within 1e-12 relative.
"""
import os
import warnings

import numpy as np
import pytest

from oracle import mesh as ORC

pytestmark = pytest.mark.gpu

# cases per counter family; HALOTOOLS_AMD_SWEEP_SEEDS=N widens the sweep
# (N count cases, N/2 marked, N/3 jackknife, N/8 medium-size)
NSEEDS = int(os.environ.get("HALOTOOLS_AMD_SWEEP_SEEDS", "24"))


@pytest.fixture(scope="module")
def mo(gpu):
    import halotools_amd.mock_observables as mo
    return mo


def _points(rng, n, box, clumped):
    if not clumped or n < 10:
        return rng.uniform(0, 1, (n, 3)) * box
    ncl = max(1, n // 50)
    cen = rng.uniform(0, 1, (ncl, 3)) * box
    p = cen[rng.randint(0, ncl, n)] + rng.normal(0, 0.01, (n, 3)) * box
    return np.mod(p, box)


def _case(seed):
    rng = np.random.RandomState(1000 + seed)
    pbc = bool(rng.randint(0, 4))  # 3 in 4 periodic
    box = np.array([1.0, 1.0, 1.0]) if rng.randint(0, 2) else rng.uniform(0.6, 1.4, 3)
    n1 = int(rng.choice([1, 7, 64, 65, 300, 1500, 4000]))
    n2 = int(rng.choice([1, 50, 129, 900, 3000]))
    clumped = bool(rng.randint(0, 2))
    s1 = _points(rng, n1, box, clumped)
    s2 = _points(rng, n2, box, clumped)
    nb = int(rng.choice([2, 3, 5, 20, 70]))
    rmax = rng.uniform(0.02, 0.3) * box.min()
    rb = np.sort(rng.uniform(0.0, rmax, nb))
    rb[-1] = rmax
    pi = np.sort(rng.uniform(0.0, rmax, int(rng.choice([2, 3, 9]))))
    pi[-1] = rmax
    cell = None if rng.randint(0, 2) else rmax * rng.uniform(0.3, 1.0)
    return s1, s2, rb, pi, (box if pbc else None), cell


def _projected_rejected(a, b, rb, pi_max, period):
    # the reference rejects pi_max >= period_z / 3, also when an unperiodic
    # sample is enclosed in a box only 3 pi_max deep (PC/npairs_projected.py:191-212)
    if period is not None:
        return pi_max >= np.atleast_1d(period)[-1] / 3.0
    box = ORC.enclose_in_box(a[:, 0], a[:, 1], a[:, 2], b[:, 0], b[:, 1], b[:, 2],
                             min_size=[rb[-1] * 3.0, rb[-1] * 3.0, pi_max * 3.0])[-1]
    return pi_max >= box[2] / 3.0


@pytest.mark.parametrize("seed", range(NSEEDS))
def test_counts_sweep(mo, seed):
    s1, s2, rb, pi, period, cell = _case(seed)
    kw = dict(period=period)
    if cell is not None:
        kw.update(approx_cell1_size=[cell] * 3, approx_cell2_size=[cell] * 3)
    for a, b in ((s1, s2), (s1, s1)):
        got = mo.npairs_3d(a, b, rb, **kw)
        assert np.array_equal(got, ORC.npairs_3d(a, b, rb, num_threads=4, **kw)), "npairs_3d"
        got = mo.npairs_xy_z(a, b, rb, pi, **kw)
        assert np.array_equal(got, ORC.npairs_xy_z(a, b, rb, pi, num_threads=4, **kw)), "npairs_xy_z"
        try:
            got = mo.npairs_projected(a, b, rb, pi[-1], **kw)
        except ValueError:
            assert _projected_rejected(a, b, rb, pi[-1], period), "projected raised"
        else:
            assert np.array_equal(got, ORC.npairs_projected(a, b, rb, pi[-1], num_threads=4, **kw)), "projected"
        got = mo.npairs_per_object_3d(a, b, rb, **kw)
        assert np.array_equal(got, ORC.npairs_per_object_3d(a, b, rb, num_threads=4, **kw)), "per_object"


@pytest.mark.parametrize("seed", range(NSEEDS // 2))
def test_marked_sweep(mo, seed):
    s1, s2, rb, pi, period, cell = _case(seed)
    rng = np.random.RandomState(2000 + seed)
    w1 = rng.uniform(0, 1, len(s1))
    w2 = rng.uniform(0, 1, len(s2))
    for a, b, wa, wb in ((s1, s2, w1, w2), (s1, s1, w1, w1)):
        got = mo.marked_npairs_3d(a, b, rb, 1, period=period, weights1=wa, weights2=wb)
        ref = ORC.marked_npairs_3d(a, b, rb, period=period, weights1=wa, weights2=wb, weight_func_id=1,
                                   num_threads=4)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("seed", range(NSEEDS // 3))
def test_marked_xy_z_sweep(mo, seed):
    s1, s2, rb, pi, period, cell = _case(seed)
    rng = np.random.RandomState(3000 + seed)
    w1 = rng.uniform(0, 1, len(s1))
    w2 = rng.uniform(0, 1, len(s2))
    for a, b, wa, wb in ((s1, s2, w1, w2), (s1, s1, w1, w1)):
        got = mo.marked_npairs_xy_z(a, b, rb, pi, period=period, weights1=wa, weights2=wb, weight_func_id=1)
        ref = ORC.marked_npairs_xy_z(a, b, rb, pi, period=period, weights1=wa, weights2=wb, weight_func_id=1,
                                     num_threads=4)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("seed", range(NSEEDS // 3))
def test_jackknife_sweep(mo, seed):
    # random tags: most waves meet several subsamples (LDS rows, then HBM)
    s1, s2, rb, pi, period, cell = _case(seed)
    rng = np.random.RandomState(4000 + seed)
    ns = int(rng.choice([2, 5, 27, 40]))
    t1 = rng.randint(1, ns + 1, len(s1))
    t2 = rng.randint(1, ns + 1, len(s2))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # samples without every tag warn, as in the reference
        for a, b, ta, tb in ((s1, s2, t1, t2), (s1, s1, t1, t1)):
            got = mo.npairs_jackknife_3d(a, b, rb, ta, tb, ns, period=period)
            ref = ORC.npairs_jackknife(a, b, rb, ta, tb, ns, period=period, num_threads=4)
            assert np.array_equal(got, ref), "jackknife 3-D"
            got = mo.npairs_jackknife_xy_z(a, b, rb, pi, ta, tb, ns, period=period)
            ref = ORC.npairs_jackknife(a, b, rb, ta, tb, ns, pi_bins=pi, period=period, num_threads=4)
            assert np.array_equal(got, ref), "jackknife xy_z"


def _medium_case(seed):
    # 20k-200k points (uniform or in tight clumps), small radii: many lane
    # groups per cell, dense histograms, several groups per count replica
    rng = np.random.RandomState(7000 + seed)
    pbc = bool(rng.randint(0, 4))
    box = np.array([1.0, 1.0, 1.0]) if rng.randint(0, 2) else rng.uniform(0.7, 1.3, 3)
    n1 = int(rng.choice([20000, 65536, 100001, 200000]))
    n2 = int(rng.choice([30000, 150000]))
    clumped = bool(rng.randint(0, 2))

    def pts(n):
        if not clumped:
            return rng.uniform(0, 1, (n, 3)) * box
        ncl = n // 40
        cen = rng.uniform(0, 1, (ncl, 3)) * box
        return np.mod(cen[rng.randint(0, ncl, n)] + rng.normal(0, 0.004, (n, 3)) * box, box)

    s1, s2 = pts(n1), pts(n2)
    rmax = rng.uniform(0.005, 0.04) * box.min()
    rb = np.sort(rng.uniform(0, rmax, int(rng.choice([2, 8, 20, 70]))))
    rb[-1] = rmax
    pi = np.sort(rng.uniform(0, rmax, int(rng.choice([2, 5, 12]))))
    pi[-1] = rmax
    w1, w2 = rng.uniform(0, 1, n1), rng.uniform(0, 1, n2)
    return s1, s2, rb, pi, (box if pbc else None), w1, w2


@pytest.mark.parametrize("seed", range(max(2, NSEEDS // 8)))
def test_medium_sweep(mo, seed):
    s1, s2, rb, pi, period, w1, w2 = _medium_case(seed)
    for a, b, wa, wb in ((s1, s2, w1, w2), (s1, s1, w1, w1)):
        assert np.array_equal(mo.npairs_3d(a, b, rb, period=period),
                              ORC.npairs_3d(a, b, rb, period=period, num_threads=8)), "npairs_3d"
        assert np.array_equal(mo.npairs_xy_z(a, b, rb, pi, period=period),
                              ORC.npairs_xy_z(a, b, rb, pi, period=period, num_threads=8)), "npairs_xy_z"
        got = mo.marked_npairs_3d(a, b, rb, 1, period=period, weights1=wa, weights2=wb)
        ref = ORC.marked_npairs_3d(a, b, rb, period=period, weights1=wa, weights2=wb, weight_func_id=1,
                                   num_threads=8)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
