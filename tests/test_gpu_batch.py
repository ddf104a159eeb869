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


def test_autotune_keeps_results(data, gpu, tmp_path, monkeypatch):
    from halotools_amd import autotune
    from halotools_amd import _engine as E
    from halotools_amd.batch import HODBatchStep, MockBatch
    monkeypatch.setenv("HALOTOOLS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_CACHE", {})
    sizes, pts, lm, logMmin, c = data
    b = MockBatch(pts, L, RB, device=gpu)
    before = b.npairs_3d().clone()
    tuned = b.tune(E.HP_MARKED_NPAIRS_3D, reps=1, targets=(2.0, 8.0), waves=(2, 8))
    assert tuned and (tmp_path / "tune.json").exists()
    assert b.tile_waves in (2, 8)
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
