"""GPU parity of ``marked_npairs_xy_z``: weighted (rp, pi) sums against the
oracle (1e-12 relative; absolute 1e-12 of the largest |sum| for marks of
mixed sign) and the reference's own known answers, which hold exactly.
The known-answer cases restate PC/test_pair_counters/test_marked_npairs_xy_z.py
(PC = halotools/mock_observables/pair_counters).
"""
import numpy as np
import pytest

from oracle import mesh as ORC
from tests.helpers import fixed_seed, random_sample

pytestmark = pytest.mark.gpu

RP = np.array([0.0, 0.1, 0.2, 0.3])
PI = np.array([0, 0.15])


@pytest.fixture(scope="module")
def mo(gpu):
    import halotools_amd.mock_observables as mo
    return mo


def _grid():
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:43-59
    g = np.linspace(0, 1 - 0.001, 10)
    xx, yy, zz = np.array(np.meshgrid(g, g, g))
    return np.vstack([xx.flatten(), yy.flatten(), zz.flatten()]).T


@pytest.mark.parametrize("period", [np.array([1.0, 1.0, 1.0]), None])
def test_product_weights_vs_oracle(mo, period):
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:62-133 (rtol 1e-5 there)
    s = random_sample(1000)
    w = np.random.RandomState(fixed_seed).random_sample((1000, 1))
    gpu = mo.marked_npairs_xy_z(s, s, RP, PI, period=period, weights1=w, weights2=w, weight_func_id=1)
    ref = ORC.marked_npairs_xy_z(s, s, RP, PI, period=period, weights1=w, weights2=w, weight_func_id=1)
    assert gpu.shape == (4, 2) and gpu.dtype == np.float64
    np.testing.assert_allclose(gpu, ref, rtol=1e-12)


@pytest.mark.parametrize("wid", list(range(1, 16)))
def test_all_wfuncs_vs_oracle(mo, wid):
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:194-248 loops every id with the right width
    from halotools_amd.mock_observables.pair_counters.marked_npairs_xy_z import _func_signature_int_from_wfunc
    width = _func_signature_int_from_wfunc(wid)
    s1, s2 = random_sample(1500, seed=wid), random_sample(1200, seed=100 + wid)
    w1 = np.random.RandomState(wid).random_sample((1500, width)) - 0.5
    w2 = np.random.RandomState(wid + 7).random_sample((1200, width)) - 0.5
    w1[:, 0] = np.round(w1[:, 0] * 4) / 4  # ties for the comparison-based marks
    w2[:, 0] = np.round(w2[:, 0] * 4) / 4
    rp = np.array([0.02, 0.05, 0.1, 0.2, 0.3])
    pi = np.array([0.0, 0.05, 0.1, 0.25])
    for per in (1.0, None):
        gpu = mo.marked_npairs_xy_z(s1, s2, rp, pi, period=per, weights1=w1, weights2=w2, weight_func_id=wid)
        ref = ORC.marked_npairs_xy_z(s1, s2, rp, pi, period=per, weights1=w1, weights2=w2, weight_func_id=wid,
                                     num_threads=4)
        np.testing.assert_allclose(gpu, ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("wid,w0,w1,factor", [
    (1, 3, None, 9.0), (2, 3, None, 6.0), (3, 2, 3, 9.0), (3, 3, 2, 4.0), (4, 2, 3, 0.0), (5, 2, 3, 0.0),
    (6, 2, 3, 0.0), (7, 1, -1, -1.0), (8, 1, 3, 3.0), (9, 1, 3, 3.0), (10, 1, 3, 0.0), (10, 1, -3, -3.0),
])
def test_uniform_weight_identities(mo, wid, w0, w1, factor):
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:251-643 (asserted with exact ==)
    pts = _grid()
    n = len(pts)
    plain = ORC.npairs_xy_z(pts, pts, RP, PI, period=np.array([1.0, 1.0, 1.0]))
    w = np.ones(n) * w0 if w1 is None else np.vstack([np.ones(n) * w0, np.ones(n) * w1]).T
    res = mo.marked_npairs_xy_z(pts, pts, RP, PI, period=np.array([1.0, 1.0, 1.0]), weights1=w, weights2=w,
                                weight_func_id=wid, approx_cell1_size=[0.3, 0.3, 0.3])
    assert np.all(res == factor * plain)


def test_ratio_weights_match_npairs_xy_z(mo):
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:646-700
    pts = _grid()
    n = len(pts)
    per = np.array([1.0, 1.0, 1.0])
    w1 = np.tile((1.0, 0.0), n).reshape((n, 2))
    plain = mo.npairs_xy_z(pts, pts, RP, PI, period=per)
    r1 = mo.marked_npairs_xy_z(pts, pts, RP, PI, period=per, weights1=w1,
                               weights2=np.ones(n * 2).reshape((n, 2)), weight_func_id=11)
    assert np.all(r1 == plain)
    r3 = mo.marked_npairs_xy_z(pts, pts, RP, PI, period=per, weights1=w1,
                               weights2=np.tile((1.0, 2.0), n).reshape((n, 2)), weight_func_id=11)
    assert np.all(r3 == 2 * plain)


def test_parallelization_arguments(mo):
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:137-191: num_threads changes nothing
    s = random_sample(1000)
    w = np.random.RandomState(fixed_seed).random_sample((1000, 1))
    a = mo.marked_npairs_xy_z(s, s, RP, PI, period=1, weights1=w, weights2=w, weight_func_id=1, num_threads=1)
    b = mo.marked_npairs_xy_z(s, s, RP, PI, period=1, weights1=w, weights2=w, weight_func_id=1, num_threads=3)
    c = mo.marked_npairs_xy_z(s, s, RP, PI, period=1, weights1=w, weights2=w, weight_func_id=1,
                              num_threads="max")
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_larger_sample_many_bins(mo):
    # 40k points with 25 x 20 (rp, pi) cells: the per-wave FP64 histogram at a realistic shape
    rng = np.random.default_rng(8)
    s1 = rng.uniform(0, 250.0, (40000, 3))
    s2 = rng.uniform(0, 250.0, (30000, 3))
    w1 = rng.uniform(0.5, 1.5, 40000)
    w2 = rng.uniform(0.5, 1.5, 30000)
    rp = np.logspace(-1, np.log10(20.0), 25)
    pi = np.linspace(1.0, 40.0, 20)
    gpu = mo.marked_npairs_xy_z(s1, s2, rp, pi, period=250.0, weights1=w1, weights2=w2, weight_func_id=1)
    ref = ORC.marked_npairs_xy_z(s1, s2, rp, pi, period=250.0, weights1=w1, weights2=w2, weight_func_id=1,
                                 num_threads=8)
    np.testing.assert_allclose(gpu, ref, rtol=1e-12)
    # unit weights reduce to npairs_xy_z
    ones1, ones2 = np.ones(40000), np.ones(30000)
    u = mo.marked_npairs_xy_z(s1, s2, rp, pi, period=250.0, weights1=ones1, weights2=ones2, weight_func_id=1)
    assert np.array_equal(u, mo.npairs_xy_z(s1, s2, rp, pi, period=250.0).astype(np.float64))


def test_device_tensor_in_out_and_repeats(mo, gpu):
    import torch
    s = random_sample(3000, seed=9)
    w = np.random.RandomState(9).random_sample(3000)
    st = torch.tensor(s, device=gpu)
    wt = torch.tensor(w, device=gpu)
    out = mo.marked_npairs_xy_z(st, st, RP, PI, period=1.0, weights1=wt, weights2=wt, weight_func_id=1)
    assert isinstance(out, torch.Tensor) and out.device == gpu and tuple(out.shape) == (4, 2)
    ref = ORC.marked_npairs_xy_z(s, s, RP, PI, period=1.0, weights1=w, weights2=w, weight_func_id=1)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-12)
    again = mo.marked_npairs_xy_z(s, s, RP, PI, period=1.0, weights1=w, weights2=w, weight_func_id=1)
    assert np.array_equal(again, out.cpu().numpy())


def test_single_point_sample(mo):
    s = random_sample(500, seed=3)
    w = np.ones(500)
    res = mo.marked_npairs_xy_z(s[:1], s, RP, PI, period=1.0, weights1=w[:1], weights2=w, weight_func_id=1)
    ref = ORC.marked_npairs_xy_z(s[:1], s, RP, PI, period=1.0, weights1=w[:1], weights2=w, weight_func_id=1)
    np.testing.assert_allclose(res, ref, rtol=1e-12)
