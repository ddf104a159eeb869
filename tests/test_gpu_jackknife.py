"""GPU parity of npairs_jackknife_3d / npairs_jackknife_xy_z against the oracle.

With unit weights (the reference's default, and what its jackknife
correlation functions use) every entry is a sum of 1 and 1/2 terms, so the
GPU tables must equal the oracle's exactly; with weights they agree to
1e-12 of row 0.  The cases cover the kernel's LDS tag rows, the HBM fallback
once a wave meets more tags than it has rows (random tags, many subsamples),
symmetric auto-counts, cross-counts and non-periodic boxes, plus the reference's
own checks (PC/test_pair_counters/test_npairs_jackknife_3d.py:55-117).
"""
import numpy as np
import pytest

from oracle import mesh as ORC
from tests.helpers import fixed_seed, random_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pc(gpu):
    from halotools_amd.mock_observables import pair_counters as pc
    return pc


def _spatial_tags(s, n, L=1.0):
    i = np.minimum((s / L * n).astype(int), n - 1)
    return 1 + (i[:, 0] * n + i[:, 1]) * n + i[:, 2]


@pytest.mark.parametrize("n", [2, 5])
@pytest.mark.parametrize("period", [1.0, None])
def test_auto_spatial_tags_exact(pc, n, period):
    s = random_sample(3000, seed=fixed_seed)
    t = _spatial_tags(s, n)
    rb = np.array([0.0, 0.02, 0.05, 0.1, 0.2, 0.3])
    got = pc.npairs_jackknife_3d(s, s, rb, t, t, n ** 3, period=period)
    ref = ORC.npairs_jackknife(s, s, rb, t, t, n ** 3, period=period, num_threads=8)
    assert got.shape == (n ** 3 + 1, len(rb))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("nsamp", [10, 40, 300])
def test_random_tags_overflow_rows_exact(pc, nsamp):
    # random tags: every wave meets many tags, most shares go to the HBM table
    s1, s2 = random_sample(2500, seed=1), random_sample(2000, seed=2)
    rng = np.random.RandomState(nsamp)
    t1 = rng.randint(1, nsamp + 1, 2500)
    t2 = rng.randint(1, nsamp + 1, 2000)
    rb = np.array([0.01, 0.05, 0.1, 0.2])
    for a, b, ta, tb in ((s1, s1, t1, t1), (s1, s2, t1, t2)):
        got = pc.npairs_jackknife_3d(a, b, rb, ta, tb, nsamp, period=1.0)
        ref = ORC.npairs_jackknife(a, b, rb, ta, tb, nsamp, period=1.0, num_threads=8)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("period", [1.0, None])
def test_weighted_cross_and_auto(pc, period):
    s1, s2 = random_sample(2000, seed=3), random_sample(1800, seed=4)
    t1, t2 = _spatial_tags(s1, 3), _spatial_tags(s2, 3)
    w1 = np.random.RandomState(5).random_sample(2000)
    w2 = np.random.RandomState(6).random_sample(1800)
    rb = np.array([0.02, 0.05, 0.1, 0.2, 0.3])
    for a, b, ta, tb, wa, wb in ((s1, s2, t1, t2, w1, w2), (s1, s1, t1, t1, w1, w1)):
        got = pc.npairs_jackknife_3d(a, b, rb, ta, tb, 27, period=period, weights1=wa, weights2=wb)
        ref = ORC.npairs_jackknife(a, b, rb, ta, tb, 27, period=period, weights1=wa, weights2=wb,
                                   num_threads=8)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref[0]).max())


def test_reference_grid_rows_identical(pc):
    # PC/test_pair_counters/test_npairs_jackknife_3d.py:55-87: on a regular grid
    # every subvolume's row is the same
    g = np.linspace(0, 1 - 0.001, 10)
    xx, yy, zz = np.array(np.meshgrid(g, g, g))
    pts = np.vstack([xx.flatten(), yy.flatten(), zz.flatten()]).T
    ix = np.floor(g / 0.5).astype(int)
    ixx, iyy, izz = np.array(np.meshgrid(ix, ix, ix))
    tags = np.ravel_multi_index([ixx.flatten(), iyy.flatten(), izz.flatten()], [2, 2, 2]) + 1
    rb = np.array([0.0, 0.1, 0.2, 0.3])
    res = pc.npairs_jackknife_3d(pts, pts, rb, tags, tags, 8, period=np.array([1.0, 1.0, 1.0]))
    for k in range(1, 7):
        assert np.all(res[k] == res[k + 1])
    s = random_sample(1000)
    with np.errstate(all="ignore"):
        jt = np.sort(np.random.RandomState(fixed_seed).randint(1, 11, size=1000))
    w = np.random.RandomState(2).random_sample(1000)
    r = pc.npairs_jackknife_3d(s, s, rb, jt, jt, 10, period=1.0, weights1=w, weights2=w)
    assert r.shape == (11, 4)
    ref = ORC.npairs_jackknife(s, s, rb, jt, jt, 10, period=1.0, weights1=w, weights2=w)
    np.testing.assert_allclose(r, ref, rtol=0, atol=1e-12 * np.abs(ref[0]).max())


@pytest.mark.parametrize("pi", [np.array([0.0, 0.2]), np.linspace(0.0, 0.3, 7)])
def test_xy_z_exact(pc, pi):
    s1, s2 = random_sample(2500, seed=7), random_sample(2000, seed=8)
    t1, t2 = _spatial_tags(s1, 4), _spatial_tags(s2, 4)
    rp = np.array([0.01, 0.05, 0.1, 0.2])
    for a, b, ta, tb in ((s1, s1, t1, t1), (s1, s2, t1, t2)):
        got = pc.npairs_jackknife_xy_z(a, b, rp, pi, ta, tb, 64, period=1.0)
        ref = ORC.npairs_jackknife(a, b, rp, ta, tb, 64, pi_bins=pi, period=1.0, num_threads=8)
        assert got.shape == (65, len(rp), len(pi))
        assert np.array_equal(got, ref)


def test_device_tensors_and_larger_box(pc, gpu):
    import torch
    rng = np.random.default_rng(9)
    s = rng.uniform(0, 250.0, (100000, 3))
    t = _spatial_tags(s, 5, L=250.0)
    rb = np.logspace(-1, np.log10(20.0), 12)
    got = pc.npairs_jackknife_3d(torch.tensor(s, device=gpu), torch.tensor(s, device=gpu), rb,
                                 torch.tensor(t, device=gpu), torch.tensor(t, device=gpu), 125, period=250.0)
    assert isinstance(got, torch.Tensor) and tuple(got.shape) == (126, 12)
    ref = ORC.npairs_jackknife(s, s, rb, t, t, 125, period=250.0, num_threads=8)
    assert np.array_equal(got.cpu().numpy(), ref)


def test_kernel_rejects_bad_tags_without_faulting(pc, gpu):
    # the library's own tag check (status bit 2): out-of-range tags are never
    # used as table indices
    import torch
    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    s = torch.tensor(random_sample(500, seed=3), device=gpu)
    cols = tuple(s[:, i].contiguous() for i in range(3))
    tags = torch.tensor(np.r_[np.full(250, 2), np.full(249, 9), [0]], dtype=torch.int32, device=gpu)
    mock = dict(n1=500, n2=500, period=[1.0] * 3, pbc=True)
    res = D.run(E.HP_NPAIRS_JACKKNIFE_3D, [mock], cols, cols, np.array([0.01, 0.04]), same=True, jtags1=tags,
                n_samples=4, device=gpu)
    torch.cuda.synchronize()
    assert int(res.cbuf[-1].item()) & 2


def test_weighted_tables_reproducible_to_rounding(pc):
    # weighted subsample rows are added with FP64 atomics from many waves: two
    # runs agree to rounding (relative to row 0), not necessarily bit for bit
    s = random_sample(3000, seed=9)
    t = _spatial_tags(s, 3)
    w = np.random.default_rng(2).uniform(0.5, 1.5, 3000)
    rb = np.array([0.02, 0.05, 0.1, 0.2])
    a = pc.npairs_jackknife_3d(s, s, rb, t, t, 27, period=1.0, weights1=w, weights2=w)
    b = pc.npairs_jackknife_3d(s, s, rb, t, t, 27, period=1.0, weights1=w, weights2=w)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-13 * np.abs(a[0]).max())
