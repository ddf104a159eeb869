"""GPU parity at the benchmark's full sizes.

The ~1M-point NFW-clustered box (npairs_3d, per-object rows, jackknife table),
2M uniform points (npairs_xy_z with 16 x 41 bins and npairs_projected), and
1M uniform halos with Zheng07 weights (product-marked sums): against the
oracle run on all host threads (bit-exact counts; marked sums within 1e-12
of the oracle's compensated sum), and
through identities that hold at any size:
  * the projected counts equal the pi = pi_max column of the xy_z counts;
  * per-object rows add up to npairs_3d;
  * the jackknife rows of N subsamples add up to (N - 1) times row 0 (every
    pair is counted at full weight in all rows but the one or two of its
    points' subsamples, where it counts half or nothing).
"""
import os

import numpy as np
import pytest

from oracle import mesh as ORC

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)


@pytest.fixture(scope="module")
def box(gpu):
    import torch
    from halotools_amd import mocks
    p = mocks.config_points(2)[0]
    return p, torch.tensor(p, device=gpu), mocks.config_rbins(2)


def test_clustered_1M_npairs_3d_exact(box):
    from halotools_amd.mock_observables import npairs_3d
    p, t, rb = box
    got = npairs_3d(t, t, rb, period=1000.0).cpu().numpy()
    assert got[-1] >= len(p)
    assert np.array_equal(got, ORC.npairs_3d(p, p, rb, period=1000.0, num_threads=THREADS))


def test_clustered_1M_per_object_rows_add_up(box):
    from halotools_amd.mock_observables import npairs_3d, npairs_per_object_3d
    p, t, rb = box
    rows = npairs_per_object_3d(t, t, rb, period=1000.0)
    assert tuple(rows.shape) == (len(p), len(rb))
    assert np.array_equal(rows.sum(0).cpu().numpy(), npairs_3d(t, t, rb, period=1000.0).cpu().numpy())
    # a sample of rows against the oracle's per-object counts of the same points
    sel = np.random.default_rng(1).choice(len(p), 300, replace=False)
    ref = ORC.npairs_per_object_3d(p[sel], p, rb, period=1000.0, num_threads=THREADS)
    assert np.array_equal(rows.cpu().numpy()[sel], ref)


def test_clustered_1M_jackknife_rows_identity(box):
    import torch
    from halotools_amd.mock_observables import npairs_3d, npairs_jackknife_3d
    p, t, rb = box
    sub = np.minimum((p / 1000.0 * 5).astype(np.int64), 4)
    tags = torch.tensor(1 + (sub[:, 0] * 5 + sub[:, 1]) * 5 + sub[:, 2], device=t.device)
    jk = npairs_jackknife_3d(t, t, rb, tags, tags, 125, period=1000.0).cpu().numpy()
    assert np.array_equal(jk[0], npairs_3d(t, t, rb, period=1000.0).cpu().numpy().astype(float))
    assert np.array_equal(jk[1:].sum(0), 124.0 * jk[0])
    assert np.all(jk[1:] <= jk[0])


def test_uniform_2M_xy_z_and_projected_exact(gpu):
    import torch
    from halotools_amd import mocks
    from halotools_amd.mock_observables import npairs_projected, npairs_xy_z
    p = mocks.config_points(3)[0]
    t = torch.tensor(p, device=gpu)
    rp, pi = mocks.config_rbins(3), np.arange(0, 41, 1.0)
    xyz = npairs_xy_z(t, t, rp, pi, period=1000.0).cpu().numpy()
    prj = npairs_projected(t, t, rp, 40.0, period=1000.0).cpu().numpy()
    assert np.array_equal(prj, xyz[:, -1])
    assert np.array_equal(xyz, ORC.npairs_xy_z(p, p, rp, pi, period=1000.0, num_threads=THREADS))


def test_uniform_1M_marked_product(gpu):
    import torch
    from halotools_amd import mocks
    from halotools_amd.mock_observables import marked_npairs_3d
    rng = np.random.default_rng([43, 4])
    p = mocks.uniform_box(1000000, 1000.0, rng)
    logm = torch.tensor(rng.uniform(11, 15, len(p)), device=gpu)
    w = mocks.zheng07_ncen(logm, torch.tensor(12.0, dtype=torch.float64, device=gpu),
                           torch.tensor(0.3, dtype=torch.float64, device=gpu))
    rb = mocks.config_rbins(4)
    t = torch.tensor(p, device=gpu)
    got = marked_npairs_3d(t, t, rb, 1, period=1000.0, weights1=w, weights2=w).cpu().numpy()
    wh = w.cpu().numpy()
    # ~1e9 terms per bin: any two summation orders (the reference's serial one
    # among them) differ by a few 1e-12, so the reference value is the
    # oracle's compensated sum (pinned against math.fsum in test_oracle.py)
    ORC.set_compensated(True)
    try:
        ref = ORC.marked_npairs_3d(p, p, rb, period=1000.0, weights1=wh, weights2=wh, weight_func_id=1,
                                   num_threads=THREADS)
    finally:
        ORC.set_compensated(False)
    np.testing.assert_allclose(got, ref, rtol=1e-12)
