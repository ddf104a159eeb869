"""CPU tests of the two-point caller layer (no GPU needed).

* The oracle (oracle/clustering.py) is pinned by closed forms: randoms equal
  to the data make every estimator exactly zero; the regular grid's known
  pair counts with analytic randoms give xi from 4/3 pi r^3 shells (and
  pi r^2 h cylinders); constant marks give M = c^2 (or 1); the shuffled
  "random" marks are the reference's aliased in-place shuffle.
* Argument processing, messages and the estimator helpers follow the
  reference's tests (TPC/tests/test_tpcf.py:456-833, test_tpcf_estimators.py,
  test_clustering_helpers.py, test_marked_tpcf.py:141-310;
  TPC = halotools/mock_observables/two_point_clustering).
"""
import warnings

import numpy as np
import pytest
import torch

from halotools_amd.custom_exceptions import HalotoolsError
from halotools_amd.mock_observables import marked_tpcf, rp_pi_tpcf, tpcf, wp
from halotools_amd.mock_observables.two_point_clustering import clustering_helpers as CH
from halotools_amd.mock_observables.two_point_clustering import tpcf_estimators as TE
from halotools_amd.mock_observables.two_point_clustering.marked_tpcf import _median0, _numpy_rng_context
from oracle import clustering as OC
from oracle import mesh as ORC
from tests.helpers import fixed_seed, generate_3d_regular_mesh, grid_rbins, random_sample

ESTIMATORS = ["Natural", "Davis-Peebles", "Hewett", "Hamilton", "Landy-Szalay"]


# ------------------------------------------------------------------ oracle pins
@pytest.mark.parametrize("est", ESTIMATORS)
def test_oracle_randoms_equal_data_give_zero(est):
    # DD = DR = RR and N = NR: every estimator is exactly 0 (a dropped term, a
    # wrong sign or factor of any of them breaks this)
    s = random_sample(300, seed=3)
    rb = np.linspace(0.02, 0.25, 6)
    xi = OC.tpcf(s, rb, randoms=s, period=np.array([1.0, 1.0, 1.0]), estimator_name=est)
    assert np.array_equal(xi, np.zeros(5))
    if est in ("Natural", "Hamilton", "Landy-Szalay"):
        s2 = random_sample(300, seed=4)
        # cross with sample2 = randoms = sample1 in a different order is still DD = DR = RR
        xi11, xi12, xi22 = OC.tpcf(s, rb, sample2=s[::-1].copy(), randoms=s, period=np.array([1.0] * 3),
                                   estimator_name=est)
        assert np.array_equal(xi12, np.zeros(5))
        assert s2.shape == s.shape


def test_oracle_grid_tpcf_closed_form():
    # 10^3 grid: cumulative counts [1, 7, 19, 27] x 1000 (TPC/../pair_counters tests)
    g = generate_3d_regular_mesh(10)
    rb = grid_rbins()
    xi = OC.tpcf(g, rb, period=np.array([1.0, 1.0, 1.0]))
    DD = np.diff([1000.0, 7000.0, 19000.0, 27000.0])
    shells = 4.0 / 3.0 * np.pi * np.diff(rb ** 3)
    np.testing.assert_allclose(xi, DD / (1000.0 ** 2 * shells) - 1.0, rtol=1e-13)


def test_oracle_grid_rp_pi_and_wp_closed_form():
    g = generate_3d_regular_mesh(10)
    rb = grid_rbins()
    pi = np.array([0.0, 0.11])
    xi = OC.rp_pi_tpcf(g, rb, pi, period=np.array([1.0, 1.0, 1.0]))
    # npairs_xy_z of the grid: rows [1,3],[5,15],[9,27],[9,27] x 1000
    DD = np.array([[8000.0], [8000.0], [0.0]])
    cyl = np.pi * np.diff(rb ** 2)[:, None] * (2 * 0.11)
    np.testing.assert_allclose(xi, DD / (1000.0 ** 2 * cyl) - 1.0, rtol=1e-13)
    w = OC.wp(g, rb, 0.11, period=np.array([1.0, 1.0, 1.0]))
    np.testing.assert_allclose(w, 2 * 0.11 * (DD[:, 0] / (1000.0 ** 2 * cyl[:, 0]) - 1.0), rtol=1e-13)


def test_oracle_marked_tpcf_constant_marks():
    s = random_sample(400, seed=5)
    rb = np.linspace(0.05, 0.25, 5)
    m = np.full(400, 3.0)
    M = OC.marked_tpcf(s, rb, marks1=m, period=np.array([1.0] * 3), normalize_by="number_counts")
    assert np.array_equal(M, np.full(4, 9.0))
    M = OC.marked_tpcf(s, rb, marks1=m, period=np.array([1.0] * 3), seed=fixed_seed)
    assert np.array_equal(M, np.ones(4))


def test_oracle_marked_tpcf_shuffle_is_in_place():
    # the reference's shuffled marks alias the marks: R11 = W(m[p], m[p]), not W(m, m[p])
    s = random_sample(500, seed=6)
    rb = np.linspace(0.05, 0.25, 5)
    m = np.random.RandomState(9).random_sample(500)
    per = np.array([1.0] * 3)
    M = OC.marked_tpcf(s, rb, marks1=m, period=per, seed=17)
    p = np.random.RandomState(17).permutation(np.arange(500))
    W = lambda a, b: np.diff(ORC.marked_npairs_3d(s, s, rb, period=per, weights1=a, weights2=b, weight_func_id=1))
    assert np.array_equal(M, W(m, m) / W(m[p], m[p]))
    assert not np.allclose(M, W(m, m) / W(m, m[p]), rtol=1e-9)


# ------------------------------------------------------------------ helpers and messages
def test_estimator_helpers():
    # TPC/tests/test_tpcf_estimators.py
    DD = np.arange(10) + 10
    for est in TE._list_estimators():
        TE._test_for_zero_division(DD, DD, DD, 100, 100, 100, 100, est)
    RR = DD.copy()
    RR[0] = 0
    for est in TE._list_estimators():
        with pytest.raises(ValueError) as err:
            TE._test_for_zero_division(DD, RR, RR, 100, 100, 100, 100, est)
        assert "you will have at least one NaN returned value" in err.value.args[0]
    assert TE._TP_estimator_requirements("Davis-Peebles") == (True, True, False)
    assert TE._TP_estimator_requirements("Hamilton") == (True, True, True)
    with pytest.raises(HalotoolsError) as err:
        TE._TP_estimator_requirements("Ron Perlman")
    assert "Input `estimator` must be one of the following" in err.value.args[0]


@pytest.mark.parametrize("est", ESTIMATORS)
def test_estimators_torch_equal_numpy(est):
    # the device estimator (torch, float64 counts) equals the NumPy one bit for bit
    rng = np.random.default_rng(1)
    DD = rng.integers(1, 10 ** 7, 12)
    DR = rng.integers(1, 10 ** 7, 12)
    RR = rng.integers(1, 10 ** 7, 12)
    N1, NR = 123457, 987651
    a = TE._TP_estimator(DD, DR, RR, N1, N1, NR, NR, est)
    t = TE._TP_estimator(*[torch.tensor(x, dtype=torch.float64) for x in (DD, DR, RR)], N1, N1, NR, NR, est)
    assert np.array_equal(a, t.numpy())
    b = OC.estimator(DD, DR, RR, N1, N1, NR, NR, est)
    assert np.array_equal(a, b)
    if est in ("Natural", "Hamilton", "Landy-Szalay"):
        D2R = rng.integers(1, 10 ** 7, 12)
        a = TE._TP_estimator_crossx(DD, DR, D2R, RR, N1, 5432, NR, NR, est)
        t = TE._TP_estimator_crossx(*[torch.tensor(x, dtype=torch.float64) for x in (DD, DR, D2R, RR)], N1, 5432,
                                    NR, NR, est)
        assert np.array_equal(a, t.numpy())
        assert np.array_equal(a, OC.estimator_crossx(DD, DR, D2R, RR, N1, 5432, NR, NR, est))
    else:
        with pytest.raises(ValueError):
            TE._TP_estimator_crossx(DD, DR, DR, RR, N1, 5432, NR, NR, est)


def test_single_bin_returns_scalar():
    xi = TE._TP_estimator(np.array([5]), None, np.array([4.0]), 10, 10, 10, 10, "Natural")
    assert np.ndim(xi) == 0 and xi == 0.25


def test_clustering_helpers():
    # TPC/tests/test_clustering_helpers.py
    with pytest.raises(ValueError) as err:
        CH.verify_tpcf_estimator("Ron Perlman")
    assert "is not in the list of available estimators" in err.value.args[0]
    s1 = random_sample(10)
    s2, same, do_cross = CH.process_optional_input_sample2(s1, None, True)
    assert same and do_cross is True and s2 is s1
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s2, same, do_cross = CH.process_optional_input_sample2(s1, s1.copy(), True)
        assert "exactly the same" in str(w[-1].message)
    assert same and do_cross is False
    s2, same, do_cross = CH.process_optional_input_sample2(s1, random_sample(11), True)
    assert not same and do_cross
    with pytest.raises(TypeError):
        CH.process_optional_input_sample2(s1, np.ones((5, 2)), True)


def test_rng_context_and_median():
    np.random.seed(5)
    before = np.random.get_state()[1].copy()
    with _numpy_rng_context(43):
        p = np.random.permutation(10)
    assert np.array_equal(p, np.random.RandomState(43).permutation(10))
    assert np.array_equal(np.random.get_state()[1], before)
    x = np.random.RandomState(2).random_sample((4, 7))
    for n in (1, 2, 3, 4):
        assert np.array_equal(_median0(torch.tensor(x[:n])).numpy(), np.median(x[:n], axis=0))


def _s():
    return random_sample(100), random_sample(101, seed=fixed_seed + 1), random_sample(102, seed=fixed_seed + 2)


def test_tpcf_messages():
    # TPC/tests/test_tpcf.py:247-316, 456-535, 768-833
    s1, s2, r = _s()
    rb = np.linspace(0.001, 0.3, 5)
    with pytest.raises(ValueError) as err:
        tpcf(s1, rb, sample2=s2, randoms=None, period=None)
    assert "If no PBCs are specified, randoms must be provided." in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        tpcf(s1, rb, sample2=s2, randoms=r, period=1.0, RR_precomputed=0.3)
    assert "``RR_precomputed`` and ``NR_precomputed`` arguments, or neither\n" in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        tpcf(s1, rb, sample2=s2, randoms=r, period=1.0, RR_precomputed=rb[:-2], NR_precomputed=102)
    assert "\nLength of ``RR_precomputed`` must match length of ``rbins``\n" in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        tpcf(s1, rb, sample2=s2, randoms=r, period=1.0, RR_precomputed=rb[:-1], NR_precomputed=5)
    assert "the value of NR_precomputed must agree with the number of randoms" in err.value.args[0]
    with pytest.raises(TypeError):
        tpcf(s1, np.array([0.1, 0.3, 0.2]), period=1.0)
    with pytest.raises(ValueError) as err:
        tpcf(s1, np.array([0.1, 0.2, 0.4]), period=1.0)
    assert "does not permit you to look for pairs separated by values" in err.value.args[0]
    with pytest.raises(TypeError):
        tpcf(s1, rb, sample2=np.ones((10, 2)), period=1.0)
    with pytest.raises(ValueError) as err:
        tpcf(s1, rb, period=1.0, do_auto="yes")
    assert "`do_auto` and `do_cross` keywords must be boolean-valued." in err.value.args[0]
    with pytest.raises(ValueError) as err:
        tpcf(s1, rb, period=1.0, estimator="Ron Perlman")
    assert "is not in the list of available estimators" in err.value.args[0]
    with pytest.raises(ValueError) as err:
        tpcf(s1, rb, period=[1.0, 1.0])
    assert "Input ``period`` must be either a scalar or a 3-element sequence." in err.value.args[0]
    with pytest.raises(ValueError) as err:
        tpcf(s1, rb, period=1.0, num_threads="many")
    assert "Input ``num_threads`` must be an integer" in err.value.args[0]


def test_rp_pi_and_wp_messages():
    s1, s2, r = _s()
    rp = np.linspace(0.001, 0.3, 5)
    with pytest.raises(ValueError) as err:
        rp_pi_tpcf(s1, rp, np.array([0.0, 0.1]), sample2=s2)
    assert "If no PBCs are specified, randoms must be provided." in err.value.args[0]
    with pytest.raises(TypeError):
        rp_pi_tpcf(s1, rp, np.array([0.0, 0.2, 0.1]), period=1.0)
    with pytest.raises(ValueError) as err:
        rp_pi_tpcf(s1, rp, np.array([0.0, 0.4]), period=1.0)
    assert "does not permit you to look for pairs separated by values" in err.value.args[0]
    with pytest.raises(ValueError) as err:
        wp(s1, rp, 0.4, period=1.0)
    assert "does not permit you to look for pairs separated by values" in err.value.args[0]
    with pytest.raises(ValueError):
        wp(s1, rp, 0.1, period=1.0, estimator="Ron Perlman")


def test_marked_tpcf_messages():
    # TPC/tests/test_marked_tpcf.py:141-310
    s1, s2, _ = _s()
    rb = np.linspace(0.001, 0.25, 5)
    w1 = np.random.RandomState(fixed_seed).random_sample(100)
    w2 = np.random.RandomState(fixed_seed).random_sample(101)
    with pytest.raises(ValueError) as err:
        marked_tpcf(s1, rb, marks1=w1, period=1, weight_func_id="a")
    assert "`wfunc` parameter must be an integer ID of the desired function." in err.value.args[0]
    with pytest.raises(ValueError) as err:
        marked_tpcf(s1, rb, marks1=w1, period=1, normalize_by="Ron Perlman")
    assert "`normalize_by` parameter not recognized." in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_tpcf(s1, rb, marks1=np.ones((100, 1, 1)), period=1)
    assert "for the input `marks1`" in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_tpcf(s1, rb, sample2=s2, marks1=w1, marks2=np.ones((101, 1, 1)), period=1)
    assert "for the input `marks2`" in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_tpcf(s1, rb, marks1=w1[:-1], period=1)
    assert "`marks1` must have same length as `sample1`." in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_tpcf(s1, rb, sample2=s2, marks1=w1, marks2=w2[:-1], period=1)
    assert "`marks2` must have same length as `sample2`." in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_tpcf(s1, rb, marks1=w1, period=1, randomize_marks=[True, False])
    assert "`randomize_marks` must have same length" in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_tpcf(s1, rb, marks1=w1, period=1, randomize_marks=np.ones((1, 1), dtype=bool))
    assert "`randomize_marks` must be one dimensional." in err.value.args[0]
    with pytest.raises(ValueError):
        marked_tpcf(s1, np.array([0.1, 0.5]), marks1=w1, period=1)
