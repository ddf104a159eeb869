"""GPU parity with more bins than one pass's LDS histogram holds.

The reference allocates its outputs at any size
(PC/cpairs/npairs_3d_engine.pyx:56, PC/cpairs/npairs_xy_z_engine.pyx:63,
PC/marked_cpairs/marked_npairs_3d_engine.pyx:76,
PC/marked_cpairs/marked_npairs_xy_z_engine.pyx:81; PC =
halotools/mock_observables/pair_counters).  Here a launch histograms as many
radial / rp start bins as its per-wave LDS area holds; wider histograms are
counted by consecutive launches over ranges of start bins (each evaluates
every pair).  Every case below is checked against the oracle: counts
bit-exact, weighted sums to 1e-12 relative; edge tables with more than one
edge per lookup bucket (thousands of edges) are covered too.
"""
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


def _edges(n, lo, hi):
    return np.concatenate([[0.0], np.logspace(np.log10(lo), np.log10(hi), n - 1)])


@pytest.mark.parametrize("nbins", [700, 5000, 60000])
@pytest.mark.parametrize("auto", [True, False])
def test_npairs_3d_many_bins(mo, nbins, auto):
    s1 = random_sample(3000, seed=11)
    s2 = s1 if auto else random_sample(2500, seed=12)
    rb = _edges(nbins, 1e-3, 0.12)
    got = mo.npairs_3d(s1, s2, rb, period=1.0)
    assert np.array_equal(got, ORC.npairs_3d(s1, s2, rb, period=1.0, num_threads=8))


@pytest.mark.parametrize("nrp,npi", [(200, 100), (450, 120)])
def test_npairs_xy_z_many_cells(mo, nrp, npi):
    s1, s2 = random_sample(3000, seed=13), random_sample(2800, seed=14)
    rp = _edges(nrp, 1e-3, 0.1)
    pi = np.linspace(0.0, 0.2, npi)
    for a, b in ((s1, s1), (s1, s2)):
        got = mo.npairs_xy_z(a, b, rp, pi, period=1.0)
        assert got.shape == (nrp, npi)
        assert np.array_equal(got, ORC.npairs_xy_z(a, b, rp, pi, period=1.0, num_threads=8))
        prj = mo.npairs_projected(a, b, rp, 0.2, period=1.0)
        assert np.array_equal(prj, got[:, -1])


@pytest.mark.parametrize("nbins,wid", [(300, 1), (300, 12), (2000, 1), (2000, 4)])
def test_marked_npairs_3d_many_bins(mo, nbins, wid):
    s1, s2 = random_sample(3000, seed=15), random_sample(2600, seed=16)
    nw = {1: 1, 4: 2, 12: 4}[wid]
    rng = np.random.default_rng(wid)
    w1, w2 = rng.uniform(0.5, 1.5, (3000, nw)), rng.uniform(0.5, 1.5, (2600, nw))
    if wid == 4:
        w1[:, 0], w2[:, 0] = np.round(w1[:, 0]), np.round(w2[:, 0])
    rb = _edges(nbins, 1e-3, 0.1)
    for a, b, wa, wb in ((s1, s1, w1, w1), (s1, s2, w1, w2)):
        got = mo.marked_npairs_3d(a, b, rb, wid, period=1.0, weights1=wa, weights2=wb)
        ref = ORC.marked_npairs_3d(a, b, rb, period=1.0, weights1=wa, weights2=wb, weight_func_id=wid,
                                   num_threads=8)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * abs(ref[-1]))


@pytest.mark.parametrize("nrp,npi", [(80, 80), (200, 100)])
def test_marked_npairs_xy_z_many_cells(mo, nrp, npi):
    s1, s2 = random_sample(2500, seed=17), random_sample(2200, seed=18)
    rng = np.random.default_rng(5)
    w1, w2 = rng.uniform(0.5, 1.5, 2500), rng.uniform(0.5, 1.5, 2200)
    rp = _edges(nrp, 1e-3, 0.1)
    pi = np.linspace(0.0, 0.2, npi)
    for a, b, wa, wb in ((s1, s1, w1, w1), (s1, s2, w1, w2)):
        got = mo.marked_npairs_xy_z(a, b, rp, pi, period=1.0, weights1=wa, weights2=wb, weight_func_id=1)
        ref = ORC.marked_npairs_xy_z(a, b, rp, pi, period=1.0, weights1=wa, weights2=wb, weight_func_id=1,
                                     num_threads=8)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * abs(ref[-1, -1]))


def test_per_object_many_bins(mo):
    s1, s2 = random_sample(2000, seed=19), random_sample(3000, seed=20)
    rb = _edges(700, 1e-3, 0.1)
    for a, b in ((s1, s1), (s1, s2)):
        got = mo.npairs_per_object_3d(a, b, rb, period=1.0)
        assert np.array_equal(got, ORC.npairs_per_object_3d(a, b, rb, period=1.0, num_threads=8))


@pytest.mark.parametrize("nrp,npi", [(0, 0), (100, 60), (400, 60)])
def test_jackknife_many_bins(mo, nrp, npi):
    from halotools_amd.mock_observables import npairs_jackknife_3d, npairs_jackknife_xy_z
    s1, s2 = random_sample(2500, seed=21), random_sample(2000, seed=22)
    t1 = 1 + np.minimum((s1[:, 0] * 3).astype(int), 2)
    t2 = 1 + np.minimum((s2[:, 0] * 3).astype(int), 2)
    for a, b, ta, tb in ((s1, s1, t1, t1), (s1, s2, t1, t2)):
        if nrp == 0:
            rb = _edges(5000, 1e-3, 0.1)
            got = npairs_jackknife_3d(a, b, rb, ta, tb, 3, period=1.0)
            ref = ORC.npairs_jackknife(a, b, rb, ta, tb, 3, period=1.0, num_threads=8)
        else:
            rp, pi = _edges(nrp, 1e-3, 0.1), np.linspace(0.0, 0.2, npi)
            got = npairs_jackknife_xy_z(a, b, rp, pi, ta, tb, 3, period=1.0)
            ref = ORC.npairs_jackknife(a, b, rp, ta, tb, 3, pi_bins=pi, period=1.0, num_threads=8)
        assert np.array_equal(got, ref)


def test_marked_product_gradient_many_bins(gpu):
    # Jacobian rows of 300 bins: counted over several launches of bin ranges
    from halotools_amd.autodiff import marked_npairs_3d_product
    s = random_sample(1500, seed=23)
    rb = _edges(300, 1e-3, 0.08)
    w = torch.tensor(np.random.default_rng(3).uniform(0.5, 1.5, 1500), device=gpu, requires_grad=True)
    c = torch.tensor(np.random.default_rng(4).normal(size=300), device=gpu)
    W = marked_npairs_3d_product(s, s, rb, w, w, period=1.0)
    (W * c).sum().backward()
    wh = w.detach().cpu().numpy()
    Wref = ORC.marked_npairs_3d(s, s, rb, period=1.0, weights1=wh, weights2=wh, weight_func_id=1)
    np.testing.assert_allclose(W.detach().cpu().numpy(), Wref, rtol=1e-12, atol=1e-12 * abs(Wref[-1]))
    # dL/dw_i = 2 sum_j w_j sum_{k: d_ij <= r_k} c_k (both factors of the auto-count)
    d = np.abs(s[:, None, :] - s[None, :, :])
    d = np.minimum(d, 1.0 - d)
    d2 = (d * d).sum(-1)
    cc = c.cpu().numpy()
    g = np.zeros(1500)
    for k, r in enumerate(rb):
        g += cc[k] * ((d2 <= r * r) * wh[None, :]).sum(axis=1)
    np.testing.assert_allclose(w.grad.cpu().numpy(), 2 * g, rtol=1e-9, atol=1e-9 * np.abs(g).max())


def test_batched_step_many_bins(gpu):
    # tangent pass (W and dW/dtheta rows plus counts) and reverse pass with 400 bins
    from halotools_amd import batch
    rng = np.random.default_rng(9)
    sizes = [3000, 800, 2000]
    L = 100.0
    pts = [torch.tensor(rng.uniform(0, L, (n, 3)), device=gpu) for n in sizes]
    lm = torch.tensor(rng.uniform(11, 15, sum(sizes)), device=gpu)
    rb = _edges(400, 0.05, 12.0)
    c = torch.tensor(rng.normal(size=400), device=gpu)
    b = batch.MockBatch(pts, L, rb, device=gpu)
    logMmin = torch.tensor([11.8, 12.1, 12.3], device=gpu)
    tan = [x.clone() for x in batch.HODBatchStep(b, lm, logMmin, 0.3, c).run()]
    rev = [x.clone() for x in batch.HODBatchStep(b, lm, logMmin, 0.3, c, mode="reverse").run()]
    assert torch.equal(tan[0], rev[0])
    for m, p in enumerate(pts):
        ph = p.cpu().numpy()
        assert np.array_equal(tan[0][m].cpu().numpy(), ORC.npairs_3d(ph, ph, rb, period=L, num_threads=8))
    np.testing.assert_allclose(tan[1].cpu().numpy(), rev[1].cpu().numpy(), rtol=1e-12)
    np.testing.assert_allclose(tan[3].cpu().numpy(), rev[3].cpu().numpy(), rtol=1e-10)


def test_shared_histogram_flush_path(mo):
    """The per-wave u32 histograms (more than 64 radial bins, and every (rp, pi)
    histogram) are flushed into 64-bit accumulators before a bin could wrap.
    A lowered flush threshold (HALOPAIR_SHARED_FLUSH_AT, read at every launch)
    makes every group flush many times; the counts must not change."""
    import os
    s1 = random_sample(4000, seed=31)
    rb = _edges(100, 1e-3, 0.1)
    rp, pi = _edges(12, 1e-3, 0.1), np.linspace(0.0, 0.2, 9)
    ref3 = ORC.npairs_3d(s1, s1, rb, period=1.0, num_threads=8)
    refx = ORC.npairs_xy_z(s1, s1, rp, pi, period=1.0, num_threads=8)
    os.environ["HALOPAIR_SHARED_FLUSH_AT"] = "20000"
    try:
        got3 = mo.npairs_3d(s1, s1, rb, period=1.0)
        gotx = mo.npairs_xy_z(s1, s1, rp, pi, period=1.0)
    finally:
        del os.environ["HALOPAIR_SHARED_FLUSH_AT"]
    assert np.array_equal(got3, ref3)
    assert np.array_equal(gotx, refx)
