"""GPU tests of the differentiable product-weighted marked sums.

Forward: equal (1e-12 relative) to the oracle's marked_npairs_3d with
weight_func_id=1.  Backward: the Jacobian rows equal a NumPy brute force; the
gradient of L = sum_k c_k W_k through Zheng07 weights agrees with central
finite differences to 1e-6 relative and with an FP64 NumPy reference gradient.
"""
import math

import numpy as np
import pytest
import torch

from oracle import brute
from oracle import mesh as ORC
from tests.helpers import random_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ad(gpu):
    from halotools_amd import autodiff
    return autodiff


def _pair_matrix(s1, s2, L):
    """dsq between every (i, j) computed like the engines: (x1 - shift) - x2 with the minimum image."""
    d = s1[:, None, :] - s2[None, :, :]
    if L is not None:
        d = np.where(d > L / 2, d - L, d)
        d = np.where(d < -L / 2, d + L, d)
    return (d ** 2).sum(axis=2)


def test_forward_matches_oracle(ad, gpu):
    s = random_sample(2000, seed=31)
    w = np.random.default_rng(1).uniform(0, 1, 2000)
    rb = np.array([0.01, 0.05, 0.1, 0.2, 0.3])
    wt = torch.tensor(w, device=gpu)
    W = ad.marked_npairs_3d_product(s, s, rb, wt, wt, period=1.0)
    ref = ORC.marked_npairs_3d(s, s, rb, 1.0, w, w, 1)
    np.testing.assert_allclose(W.detach().cpu().numpy(), ref, rtol=1e-12)


@pytest.mark.parametrize("period", [1.0, None])
def test_cross_rows_match_brute_force(ad, gpu, period):
    from halotools_amd import _device as D, _engine as E
    s1, s2 = random_sample(700, seed=2), random_sample(600, seed=3)
    w1 = np.random.default_rng(4).uniform(0.5, 2, 700)
    w2 = np.random.default_rng(5).uniform(0.5, 2, 600)
    rb = np.array([0.05, 0.1, 0.2, 0.3])
    t1, t2 = torch.tensor(w1, device=gpu, requires_grad=True), torch.tensor(w2, device=gpu, requires_grad=True)
    W = ad.marked_npairs_3d_product(s1, s2, rb, t1, t2, period=period)
    c = torch.tensor([1.0, -2.0, 0.5, 3.0], device=gpu, dtype=torch.float64)
    (W * c).sum().backward()
    # brute-force gradient: dL/dw1_i = sum_k c_k sum_{j: d_ij <= r_k} w2_j
    if period is None:
        d2 = _pair_matrix(s1, s2, None)
    else:
        d2 = _pair_matrix(s1, s2, 1.0)
    g1 = np.zeros(700)
    g2 = np.zeros(600)
    for k, r in enumerate(rb):
        inb = d2 <= r * r
        g1 += c[k].item() * (inb * w2[None, :]).sum(axis=1)
        g2 += c[k].item() * (inb * w1[:, None]).sum(axis=0)
    np.testing.assert_allclose(t1.grad.cpu().numpy(), g1, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(t2.grad.cpu().numpy(), g2, rtol=1e-12, atol=1e-12)
    Wref = brute.wnpairs(s1, s2, rb, period=period, weights1=w1, weights2=w2)
    np.testing.assert_allclose(W.detach().cpu().numpy(), Wref, rtol=1e-12)


def test_hod_gradient_vs_finite_differences(ad, gpu):
    # Zheng07 central occupation weights on uniform halos, L = sum_k c_k W_k
    from halotools_amd import mocks
    rng = np.random.default_rng([43, 4])
    n = 8000
    pts = rng.uniform(0, 60.0, (n, 3))
    logm = torch.tensor(rng.uniform(11, 15, n), device=gpu)
    rb = np.logspace(-1, np.log10(10), 8)
    c = torch.tensor(rng.normal(size=8), device=gpu)
    pts_t = torch.tensor(pts, device=gpu)

    def loss(theta):
        w = mocks.zheng07_ncen(logm, theta[0], theta[1])
        return (ad.marked_npairs_3d_product(pts_t, pts_t, rb, w, w, period=60.0) * c).sum()

    theta = torch.tensor([12.0, 0.3], device=gpu, dtype=torch.float64, requires_grad=True)
    L0 = loss(theta)
    L0.backward()
    grad = theta.grad.detach().cpu().numpy()
    for a, h in ((0, 1e-4), (1, 1e-5)):
        tp = theta.detach().clone(); tp[a] += h
        tm = theta.detach().clone(); tm[a] -= h
        fd = (loss(tp).item() - loss(tm).item()) / (2 * h)
        assert abs(fd - grad[a]) <= 1e-6 * abs(fd), (a, fd, grad[a])

    # FP64 NumPy reference gradient: dW/dw_i over both factors, chained through erf
    w = 0.5 * (1 + np.array([math.erf(v) for v in (logm.cpu().numpy() - 12.0) / 0.3]))
    dw_dmin = -np.exp(-((logm.cpu().numpy() - 12.0) / 0.3) ** 2) / (math.sqrt(math.pi) * 0.3)
    cc = c.cpu().numpy()
    gw = np.zeros(n)
    # neighbours within the largest edge via a coarse grid (brute force per cell block)
    for i0 in range(0, n, 1000):
        blk = np.arange(i0, min(i0 + 1000, n))
        d2 = _pair_matrix(pts[blk], pts, 60.0)
        for k, r in enumerate(rb):
            gw[blk] += cc[k] * ((d2 <= r * r) * w[None, :]).sum(axis=1)
    gref = 2 * (gw * dw_dmin).sum()
    assert abs(gref - grad[0]) <= 1e-9 * abs(gref), (gref, grad[0])


def test_backward_deterministic(ad, gpu):
    s = random_sample(5000, seed=7)
    w = torch.tensor(np.random.default_rng(9).uniform(0, 1, 5000), device=gpu, requires_grad=True)
    rb = np.array([0.02, 0.05, 0.1])
    grads = []
    for _ in range(2):
        w.grad = None
        ad.marked_npairs_3d_product(s, s, rb, w, w, period=1.0).sum().backward()
        grads.append(w.grad.clone())
    assert torch.equal(grads[0], grads[1])


def _seam_pair():
    """Two points across the x seam of the unit box whose two pair orientations,
    (x1 + L) - x2 and (x2 - L) - x1, square to different doubles, and an edge
    r with r*r between them."""
    rng = np.random.default_rng(0)
    while True:
        xi, xj = rng.uniform(0.0, 0.05), rng.uniform(0.95, 1.0)
        df, dr = ((xi + 1.0) - xj) ** 2, ((xj - 1.0) - xi) ** 2
        if df != dr:
            break
    lo, hi = min(df, dr), max(df, dr)
    r = np.sqrt(lo)
    for _ in range(100):
        if lo <= r * r < hi:
            break
        r = np.nextafter(r, 0.0) if r * r >= hi else np.nextafter(r, 1.0)
    assert lo <= r * r < hi
    return xi, xj, r


def test_seam_correction_matches_swapped_pass(gpu):
    from halotools_amd import _device as D, _engine as E
    xi, xj, r = _seam_pair()
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 1, (600, 3))
    pts[10] = [xi, 0.5, 0.5]
    pts[20] = [xj, 0.5, 0.5]
    w = rng.integers(1, 4, 600).astype(np.float64)  # small integers: every sum is exact
    rb = np.array([0.5 * r, r, 0.2])
    t = torch.tensor(pts.T.copy(), device=gpu)
    cols = (t[0], t[1], t[2])
    cols_copy = tuple(c.clone() for c in cols)
    wt = torch.tensor(w, device=gpu).reshape(-1, 1)
    mock = [dict(n1=600, n2=600, period=[1.0] * 3, pbc=True)]
    elim = D.run(E.HP_MARKED_NPAIRS_3D, mock, cols, cols, rb * rb, same=True, w1=wt, nw=1, wfunc=1,
                 want_counts=False, want_sums=True, rows1=True, rows2=True, device=gpu)
    swapped = D.run(E.HP_MARKED_NPAIRS_3D, mock, cols, cols_copy, rb * rb, same=False, w1=wt, w2=wt.clone(), nw=1,
                    wfunc=1, want_counts=False, want_sums=True, rows1=True, rows2=True, device=gpu)
    torch.cuda.synchronize()
    r1e, r2e = elim.rows1.cpu().numpy(), elim.rows2.cpu().numpy()
    r1s, r2s = swapped.rows1.cpu().numpy(), swapped.rows2.cpu().numpy()
    assert np.array_equal(r1e, r1s)
    assert np.array_equal(r2e, r2s)
    # the two special points really are binned differently by the two orientations
    assert not np.array_equal(r1e[[10, 20]], r2e[[10, 20]])
    assert np.array_equal(elim.sums.cpu().numpy(), swapped.sums.cpu().numpy())


def test_seam_correction_non_integer_weights(gpu):
    # with non-integer weights rows2 = rows1 + corr equals the swapped pass to
    # the rounding of the row sums (removing w_i from a rounded sum is not exact)
    from halotools_amd import _device as D, _engine as E
    xi, xj, r = _seam_pair()
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 1, (800, 3))
    pts[10] = [xi, 0.5, 0.5]
    pts[20] = [xj, 0.5, 0.5]
    w = rng.uniform(0.1, 2.0, 800)
    rb = np.array([0.5 * r, r, 0.15, 0.2])
    t = torch.tensor(pts.T.copy(), device=gpu)
    cols = (t[0], t[1], t[2])
    cols_copy = tuple(c.clone() for c in cols)
    wt = torch.tensor(w, device=gpu).reshape(-1, 1)
    mock = [dict(n1=800, n2=800, period=[1.0] * 3, pbc=True)]
    elim = D.run(E.HP_MARKED_NPAIRS_3D, mock, cols, cols, rb * rb, same=True, w1=wt, nw=1, wfunc=1,
                 want_counts=False, want_sums=True, rows1=True, rows2=True, device=gpu)
    swapped = D.run(E.HP_MARKED_NPAIRS_3D, mock, cols, cols_copy, rb * rb, same=False, w1=wt, w2=wt.clone(), nw=1,
                    wfunc=1, want_counts=False, want_sums=True, rows1=True, rows2=True, device=gpu)
    torch.cuda.synchronize()
    r2e, r2s = elim.rows2.cpu().numpy(), swapped.rows2.cpu().numpy()
    np.testing.assert_array_equal(elim.rows1.cpu().numpy(), swapped.rows1.cpu().numpy())
    np.testing.assert_allclose(r2e, r2s, rtol=1e-13, atol=1e-13 * np.abs(r2s).max())
    np.testing.assert_allclose(elim.sums.cpu().numpy(), swapped.sums.cpu().numpy(), rtol=1e-13)
