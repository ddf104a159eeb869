"""GPU tests of the batched many-mock path and its HIP-graph capture.

Batched results must equal the one-mock-at-a-time results: counts exactly,
marked sums and gradients to 1e-12 relative.  A captured graph is replayed
several times with new parameters and eager work in between, and every replay
is checked against an eager step.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

L = 200.0
RB = np.logspace(-1, np.log10(15), 10)


@pytest.fixture(scope="module")
def data(gpu):
    rng = np.random.default_rng([43, 5])
    sizes = [3000, 17000, 500, 8000, 1, 25000]
    pts = [torch.tensor(rng.uniform(0, L, (n, 3)), device=gpu) for n in sizes]
    lm = [torch.tensor(rng.uniform(11, 15, n), device=gpu) for n in sizes]
    logMmin = torch.tensor(rng.uniform(11.5, 12.5, len(sizes)), device=gpu)
    c = torch.tensor(rng.normal(size=len(RB)), device=gpu)
    return sizes, pts, lm, logMmin, c


def test_batch_counts_equal_single(data, gpu):
    from halotools_amd.batch import MockBatch
    from halotools_amd.mock_observables import npairs_3d
    sizes, pts, lm, logMmin, c = data
    b = MockBatch(pts, L, RB, device=gpu)
    counts = b.npairs_3d().cpu().numpy()
    for m, p in enumerate(pts):
        assert np.array_equal(counts[m], npairs_3d(p, p, RB, period=L).cpu().numpy()), m


def test_batch_marked_and_grad_equal_single(data, gpu):
    from halotools_amd.autodiff import marked_npairs_3d_product
    from halotools_amd.batch import MockBatch
    sizes, pts, lm, logMmin, c = data
    b = MockBatch(pts, L, RB, device=gpu)
    w = torch.cat([0.5 + 0.1 * torch.sin(x) for x in lm]).requires_grad_(True)
    W = b.marked_npairs_3d_product(w)
    (W * c).sum().backward()
    offs = b.offsets
    for m, p in enumerate(pts):
        wm = w.detach()[offs[m]:offs[m + 1]].clone().requires_grad_(True)
        Wm = marked_npairs_3d_product(p, p, RB, wm, wm, period=L)
        (Wm * c).sum().backward()
        np.testing.assert_allclose(W[m].detach().cpu().numpy(), Wm.detach().cpu().numpy(), rtol=1e-12, atol=0)
        np.testing.assert_allclose(w.grad[offs[m]:offs[m + 1]].cpu().numpy(), wm.grad.cpu().numpy(), rtol=1e-12,
                                   atol=1e-12)


def test_hod_step_eager_and_graph(data, gpu):
    from halotools_amd.batch import HODBatchStep, MockBatch, one_at_a_time
    sizes, pts, lm, logMmin, c = data
    b = MockBatch(pts, L, RB, device=gpu)
    step = HODBatchStep(b, torch.cat(lm), logMmin, 0.3, c)
    counts, W, loss, grad = [t.clone() for t in step.run()]
    assert step.status() == 0
    ref = one_at_a_time(pts, lm, logMmin, 0.3, RB, c, L)
    for m in range(len(sizes)):
        assert np.array_equal(counts[m].cpu().numpy(), ref[0][m].cpu().numpy())
        np.testing.assert_allclose(W[m].cpu().numpy(), ref[1][m].cpu().numpy(), rtol=1e-12)
        np.testing.assert_allclose(grad[m].item(), ref[3][m].item(), rtol=1e-10, atol=1e-12)

    step.capture()
    rng = np.random.default_rng(11)
    for rep in range(3):
        new = torch.tensor(rng.uniform(11.5, 12.5, len(sizes)), device=gpu)
        step.logMmin.copy_(new)
        g_counts, g_W, g_loss, g_grad = [t.clone() for t in step.replay()]
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=gpu).sum().item()  # eager work between replays
        e_counts, e_W, e_loss, e_grad = [t.clone() for t in step.run()]
        torch.cuda.synchronize()
        assert torch.equal(g_counts, e_counts)
        assert torch.allclose(g_W, e_W, rtol=1e-14, atol=0)
        assert torch.allclose(g_grad, e_grad, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_tile_waves_do_not_change_results(data, gpu, waves):
    # the workgroup shape only regroups lane groups: counts and per-group
    # partial sums (reduced in a fixed order) are bit-identical
    from halotools_amd.batch import MockBatch
    sizes, pts, lm, logMmin, c = data
    ref = MockBatch(pts, L, RB, device=gpu, cells=[5.0] * len(sizes))
    b = MockBatch(pts, L, RB, device=gpu, cells=[5.0] * len(sizes), tile_waves=waves)
    assert torch.equal(b.npairs_3d(), ref.npairs_3d())
    w = torch.cat([0.5 + 0.1 * torch.sin(x) for x in lm])
    assert torch.equal(b.marked_npairs_3d_product(w), ref.marked_npairs_3d_product(w))


@pytest.mark.parametrize("which", ["reverse", "tangent"])
def test_autotune_keeps_results(data, gpu, tmp_path, monkeypatch, which):
    from halotools_amd import autotune
    from halotools_amd import _engine as E
    from halotools_amd.batch import HODBatchStep, MockBatch
    monkeypatch.setenv("HALOTOOLS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_CACHE", {})
    sizes, pts, lm, logMmin, c = data
    b = MockBatch(pts, L, RB, device=gpu)
    before = b.npairs_3d().clone()
    kind = E.HP_MARKED_NPAIRS_3D if which == "reverse" else autotune.TANGENT
    tuned = b.tune(kind, reps=1, targets=(2.0, 8.0), waves=(2, 8))
    assert tuned and (tmp_path / "tune.json").exists()
    assert b.waves_for(kind) in (2, 8)
    # the shape tuned for one kind is not applied to launches of another
    other = autotune.TANGENT if which == "reverse" else E.HP_MARKED_NPAIRS_3D
    assert b.waves_for(other) == 0 and b.waves_for(E.HP_NPAIRS_3D) == 0
    assert torch.equal(b.npairs_3d(), before)
    step = HODBatchStep(b, torch.cat(lm), logMmin, 0.3, c)
    counts = step.run()[0]
    assert torch.equal(counts, before)


def test_segment_sum_fixed_order(gpu):
    from halotools_amd import _device as D
    rng = np.random.default_rng(3)
    sizes = [0, 1, 4095, 4096, 4097, 100000, 0, 3]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    v = torch.tensor(rng.normal(size=int(offs[-1])), device=gpu)
    seg = D.SegmentSum(offs, gpu)
    out = torch.empty(len(sizes), dtype=torch.float64, device=gpu)
    got = seg(v, out).cpu().numpy().copy()
    host = v.cpu().numpy()
    for m in range(len(sizes)):
        ref = math.fsum(host[offs[m]:offs[m + 1]])
        assert abs(got[m] - ref) <= 1e-13 * max(1.0, np.abs(host[offs[m]:offs[m + 1]]).sum()), m
    again = seg(v, torch.empty_like(out)).cpu().numpy()
    assert np.array_equal(got, again)


@pytest.mark.parametrize("mode", ["tangent", "reverse"])
def test_hod_step_modes_agree(data, gpu, mode):
    from halotools_amd.batch import HODBatchStep, MockBatch, one_at_a_time
    sizes, pts, lm, logMmin, c = data
    b = MockBatch(pts, L, RB, device=gpu)
    counts, W, loss, grad = [t.clone() for t in HODBatchStep(b, torch.cat(lm), logMmin, 0.3, c, mode=mode).run()]
    ref = one_at_a_time(pts, lm, logMmin, 0.3, RB, c, L)
    for m in range(len(sizes)):
        assert np.array_equal(counts[m].cpu().numpy(), ref[0][m].cpu().numpy())
        np.testing.assert_allclose(W[m].cpu().numpy(), ref[1][m].cpu().numpy(), rtol=1e-12)
        np.testing.assert_allclose(grad[m].item(), ref[3][m].item(), rtol=1e-10, atol=1e-12)


def _tangent_pass(pts1, pts2, w1, t1, w2, t2, rbins, period, gpu):
    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    same = pts2 is None
    cols = lambda p: tuple(p[:, i].contiguous() for i in range(3))
    s1 = cols(pts1)
    s2 = s1 if same else cols(pts2)
    wt1 = torch.stack([w1, t1], dim=1).contiguous()
    wt2 = None if same else torch.stack([w2, t2], dim=1).contiguous()
    n1, n2 = len(pts1), len(pts1) if same else len(pts2)
    Lb = period if period is not None else 400.0
    mock = dict(n1=n1, n2=n2, period=[Lb] * 3, pbc=period is not None)
    res = D.run(E.HP_MARKED_NPAIRS_3D, [mock], s1, s2, rbins * rbins, same=same, w1=wt1, w2=wt2, nw=2, wfunc=1,
                want_counts=False, want_sums=True, also_counts=True, tangent=True, device=gpu)
    torch.cuda.synchronize()
    assert int(res.cbuf[-1].item()) & 0xFFFFFFFF == 0  # int32 status word
    return res


@pytest.mark.parametrize("same,period", [(True, 200.0), (False, 200.0), (True, None)])
def test_tangent_pass_matches_reverse_mode_and_fd(gpu, same, period):
    # W and dW/dtheta of the tangent pass against the product-weighted sums,
    # their derivative through the hand-written backward (reverse mode) and
    # central finite differences; counts against npairs_3d
    from halotools_amd.autodiff import marked_npairs_3d_product
    from halotools_amd.mock_observables import npairs_3d
    rng = np.random.default_rng([11, int(same), int(period is None)])
    n1, n2 = 6000, 4500
    p1 = torch.tensor(rng.uniform(0, 200.0, (n1, 3)), device=gpu)
    p2 = p1 if same else torch.tensor(rng.uniform(0, 200.0, (n2, 3)), device=gpu)
    # points on the periodic seam exercise the shifted images
    p1[:50, 0] = 200.0 - 1e-9 * torch.arange(50, device=gpu)
    if period is None:
        # the drop-ins shift non-periodic samples by their smallest coordinate;
        # starting at 0 makes that shift exact, so both paths see the same inputs
        m = min(float(p1.min()), float(p2.min()))
        p1 = p1 - m
        p2 = p1 if same else p2 - m
    x1 = torch.tensor(rng.uniform(-1, 1, n1), device=gpu)
    x2 = x1 if same else torch.tensor(rng.uniform(-1, 1, len(p2)), device=gpu)
    theta = 0.3
    w = lambda x, th: torch.sigmoid(x + th)
    dw = lambda x, th: torch.sigmoid(x + th) * (1 - torch.sigmoid(x + th))
    rb = np.logspace(-0.5, np.log10(15.0), 12)
    res = _tangent_pass(p1, None if same else p2, w(x1, theta), dw(x1, theta), w(x2, theta), dw(x2, theta),
                        rb, period, gpu)
    th = torch.tensor(theta, dtype=torch.float64, device=gpu, requires_grad=True)
    W = marked_npairs_3d_product(p1, p2, rb, w(x1, th), w(x2, th), period=period)
    np.testing.assert_allclose(res.sums[0].cpu().numpy(), W.detach().cpu().numpy(), rtol=1e-12)
    c = torch.tensor(rng.normal(size=len(rb)), device=gpu)
    (W * c).sum().backward()
    dtan = float((res.tangents[0] * c).sum())
    np.testing.assert_allclose(dtan, th.grad.item(), rtol=1e-10)
    h = 1e-5
    wp = marked_npairs_3d_product(p1, p2, rb, w(x1, theta + h), w(x2, theta + h), period=period)
    wm = marked_npairs_3d_product(p1, p2, rb, w(x1, theta - h), w(x2, theta - h), period=period)
    fd = ((wp - wm) / (2 * h)).cpu().numpy()
    np.testing.assert_allclose(res.tangents[0].cpu().numpy(), fd, rtol=1e-6, atol=1e-9)
    counts = npairs_3d(p1, p2, rb, period=period)
    assert torch.equal(res.counts[0], counts)


def test_zheng07_tangent_kernel_matches_torch(gpu):
    """The batched step's weights and logMmin tangents come from one library
    kernel (hp_zheng07_tangent); they must equal the elementwise torch
    expression of empirical_models/occupation_models/zheng07_components.py:149-153
    bit for bit (same operations, same order, same erf/exp)."""
    from halotools_amd import _device as D
    rng = np.random.default_rng(11)
    sizes = [7, 0, 5000, 1, 130000]
    mock_id = torch.from_numpy(np.repeat(np.arange(len(sizes)), sizes)).to(gpu)
    log_mass = torch.tensor(rng.uniform(10.0, 16.0, sum(sizes)), device=gpu)
    logMmin = torch.tensor(rng.uniform(11.5, 12.5, len(sizes)), device=gpu)
    for sigma in (0.3, 0.05, 1.7):
        out = torch.empty((sum(sizes), 2), dtype=torch.float64, device=gpu)
        D.zheng07_tangent(log_mass, mock_id, logMmin, sigma, out)
        u = (log_mass - logMmin[mock_id]) / sigma
        w = 0.5 * (1.0 + torch.erf(u))
        dw = -torch.exp(-u * u) / (math.sqrt(math.pi) * sigma)
        assert torch.equal(out[:, 0], w) and torch.equal(out[:, 1], dw), sigma
    with pytest.raises(ValueError):
        D.zheng07_tangent(log_mass, mock_id, logMmin, 0.3, torch.empty((3, 2), dtype=torch.float64, device=gpu))
