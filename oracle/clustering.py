"""CPU oracle of the two-point caller layer: NumPy restatement of the reference's
tpcf / rp_pi_tpcf / wp / marked_tpcf over the oracle's pair counts.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ -- never by the product package.

Follows halotools/mock_observables/two_point_clustering/ (TPC/):
* ``tpcf``         -- TPC/tpcf.py:39-209 (counts), :430-498 (estimator dispatch)
* ``rp_pi_tpcf``   -- TPC/rp_pi_tpcf.py:230-460
* ``wp``           -- TPC/wp.py:194-240
* ``marked_tpcf``  -- TPC/marked_tpcf.py:311-494, including the in-place
  shuffle of ``random_counts`` (the shuffled marks alias the marks, :420-428)
* ``estimator`` / ``estimator_crossx`` -- TPC/tpcf_estimators.py:12-129
  (scalar sample sizes only; the jackknife form is not restated)
Input validation is not restated: the oracle assumes valid, already-processed
arguments (period a length-3 array or None, samples (N, 3) arrays).
"""
from math import gamma

import numpy as np

from . import mesh as ORC


def estimator(DD, DR, RR, N1, N2, NR1, NR2, name):
    """TPC/tpcf_estimators.py:12-73 for scalar sample sizes."""
    f11 = N1 * N2 / (NR1 * NR2)
    if name == "Natural":
        xi = (1.0 / f11) * (DD / RR) - 1.0
    elif name == "Davis-Peebles":
        xi = (1.0 / (N1 * N2 / (N1 * NR2))) * (DD / DR) - 1.0
    elif name == "Hewett":
        xi = (1.0 / f11) * (DD / RR) - (1.0 / (N1 * NR2 / (NR1 * NR2))) * (DR / RR)
    elif name == "Hamilton":
        xi = (DD * RR) / (DR * DR) - 1.0
    elif name == "Landy-Szalay":
        xi = (1.0 / f11) * (DD / RR) - (1.0 / (N1 * NR2 / (NR1 * NR2))) * (2.0 * DR / RR) + 1.0
    return xi[0] if np.shape(xi)[0] == 1 else xi


def estimator_crossx(DD, D1R, D2R, RR, N1, N2, NR1, NR2, name):
    """TPC/tpcf_estimators.py:76-129 for scalar sample sizes."""
    f11 = N1 * N2 / (NR1 * NR2)
    f12 = N1 * NR2 / (NR1 * NR2)
    if name == "Natural":
        xi = (1.0 / f11) * (DD / RR) - 1.0
    elif name == "Hamilton":
        xi = (DD * RR) / (D1R * D2R) - 1.0
    elif name == "Landy-Szalay":
        xi = (1.0 / f11) * (DD / RR) - (1.0 / f12) * (D1R / RR) - (1.0 / f12) * (D2R / RR) + 1.0
    return xi[0] if np.shape(xi)[0] == 1 else xi


_NEEDS = {"Natural": (False, True), "Davis-Peebles": (True, False), "Hewett": (True, True),
          "Hamilton": (True, True), "Landy-Szalay": (True, True)}  # (DR, RR), TPC/clustering_helpers.py:16-23


def _same(a, b):
    return b is None or (a.shape == b.shape and np.all(a == b))


def _finish(auto_only, do_auto, do_cross, xi11, xi12, xi22):
    if auto_only:
        return xi11()
    if do_auto and do_cross:
        return xi11(), xi12(), xi22()
    if do_cross:
        return xi12()
    if do_auto:
        return xi11(), xi22()


def tpcf(sample1, rbins, sample2=None, randoms=None, period=None, do_auto=True, do_cross=True,
         estimator_name="Natural", num_threads=1):
    """TPC/tpcf.py:212-498 (analytic randoms when ``randoms`` is None)."""
    auto_only = _same(sample1, sample2)
    sample2 = sample1 if sample2 is None else sample2
    if auto_only:
        do_cross = False
    need_dr, need_rr = _NEEDS[estimator_name]
    cnt = lambda a, b: np.diff(ORC.npairs_3d(a, b, rbins, period=period, num_threads=num_threads))
    N1, N2 = len(sample1), len(sample2)
    D11 = cnt(sample1, sample1) if do_auto else None
    D12 = D11 if auto_only else (cnt(sample1, sample2) if do_cross else None)
    D22 = D11 if auto_only else (cnt(sample2, sample2) if do_auto else None)
    if randoms is not None:
        NR = len(randoms)
        RR = cnt(randoms, randoms) if need_rr else None
        D1R = cnt(sample1, randoms) if need_dr else None
        D2R = cnt(sample2, randoms) if (need_dr and not auto_only) else None
    else:
        NR = N1
        dv = np.diff((np.pi ** 1.5 / gamma(2.5)) * rbins ** 3)  # shell volumes, TPC/tpcf.py:66-71,127-128
        V = period.prod()
        D1R = NR * (dv * (N1 / V))
        D2R = NR * (dv * (N2 / V))
        RR = dv * (NR ** 2 / V)
    return _finish(auto_only, do_auto, do_cross,
                   lambda: estimator(D11, D1R, RR, N1, N1, NR, NR, estimator_name),
                   lambda: estimator_crossx(D12, D1R, D2R, RR, N1, N2, NR, NR, estimator_name),
                   lambda: estimator(D22, D2R, RR, N2, N2, NR, NR, estimator_name))


def rp_pi_tpcf(sample1, rp_bins, pi_bins, sample2=None, randoms=None, period=None, do_auto=True, do_cross=True,
               estimator_name="Natural", num_threads=1):
    """TPC/rp_pi_tpcf.py:230-460 (analytic cylinder randoms when ``randoms`` is None)."""
    auto_only = _same(sample1, sample2)
    sample2 = sample1 if sample2 is None else sample2
    if auto_only:
        do_cross = False
    need_dr, need_rr = _NEEDS[estimator_name]

    def cnt(a, b):
        c = ORC.npairs_xy_z(a, b, rp_bins, pi_bins, period=period, num_threads=num_threads)
        return np.diff(np.diff(c, axis=0), axis=1)
    N1, N2 = len(sample1), len(sample2)
    D11 = cnt(sample1, sample1)
    D12 = D11 if auto_only else (cnt(sample1, sample2) if do_cross else None)
    D22 = D11 if auto_only else (cnt(sample2, sample2) if do_auto else None)
    if randoms is not None:
        NR = len(randoms)
        RR = cnt(randoms, randoms) if need_rr else None
        D1R = cnt(sample1, randoms) if need_dr else None
        D2R = cnt(sample2, randoms) if (need_dr and not auto_only) else None
    else:
        NR = N1
        v = np.pi * np.outer(rp_bins ** 2.0, 2.0 * pi_bins)  # TPC/rp_pi_tpcf.py:353-357
        dv = np.diff(np.diff(v, axis=0), axis=1)
        V = period.prod()
        D1R = N1 * (dv * (N1 / V))
        D2R = N2 * (dv * (N2 / V))
        RR = dv * (NR ** 2 / V)
    return _finish(auto_only, do_auto, do_cross,
                   lambda: estimator(D11, D1R, RR, N1, N1, NR, NR, estimator_name),
                   lambda: estimator_crossx(D12, D1R, D2R, RR, N1, N2, NR, NR, estimator_name),
                   lambda: estimator(D22, D2R, RR, N2, N2, NR, NR, estimator_name))


def wp(sample1, rp_bins, pi_max, sample2=None, randoms=None, period=None, do_auto=True, do_cross=True,
       estimator_name="Natural"):
    """TPC/wp.py:194-240: 2 pi_max xi(rp, 0 <= pi <= pi_max)."""
    pi_max = float(pi_max)
    r = rp_pi_tpcf(sample1, rp_bins, np.array([0.0, pi_max]), sample2=sample2, randoms=randoms, period=period,
                   do_auto=do_auto, do_cross=do_cross, estimator_name=estimator_name)
    if isinstance(r, tuple):
        return tuple(2.0 * x[:, 0] * pi_max for x in r)
    return 2.0 * r[:, 0] * pi_max


def marked_tpcf(sample1, rbins, sample2=None, marks1=None, marks2=None, period=None, do_auto=True, do_cross=True,
                weight_func_id=1, normalize_by="random_marks", iterations=1, randomize_marks=None, seed=None):
    """TPC/marked_tpcf.py:311-494 (marks as (N,) or (N, nw) float arrays)."""
    auto_only = _same(sample1, sample2)
    sample2 = sample1 if sample2 is None else sample2
    if auto_only:
        do_cross = False
    m1 = np.ones((len(sample1), 1)) if marks1 is None else np.array(marks1, dtype=float).reshape(len(sample1), -1)
    m2 = np.ones((len(sample2), 1)) if marks2 is None else np.array(marks2, dtype=float).reshape(len(sample2), -1)
    rand = np.ones(m1.shape[1], dtype=bool) if randomize_marks is None else np.atleast_1d(randomize_marks)

    def W(a, b, wa, wb):
        return np.diff(ORC.marked_npairs_3d(a, b, rbins, period=period, weights1=wa, weights2=wb,
                                            weight_func_id=weight_func_id))

    def plain(a, b):
        return np.diff(ORC.npairs_3d(a, b, rbins, period=period))

    W11 = W(sample1, sample1, m1, m1) if do_auto else None
    W12 = W11 if auto_only else (W(sample1, sample2, m1, m2) if do_cross else None)
    W22 = W11 if auto_only else (W(sample2, sample2, m2, m2) if do_auto else None)

    def shuffled():
        # NumpyRNGContext(seed): seed the global generator, draw, restore its state
        state = np.random.get_state()
        np.random.seed(seed)
        p1 = np.random.permutation(np.arange(0, len(sample1)))
        p2 = np.random.permutation(np.arange(0, len(sample2)))
        np.random.set_state(state)
        # the shuffled copies ARE m1 / m2 (TPC/marked_tpcf.py:420-428)
        for i in range(m1.shape[1]):
            if rand[i]:
                m1[:, i] = m1[p1, i]
        for i in range(m2.shape[1]):
            if rand[i]:
                m2[:, i] = m2[p2, i]
        r11 = W(sample1, sample1, m1, m1) if do_auto else None
        r12 = r11 if auto_only else (W(sample1, sample2, m1, m2) if do_cross else None)
        r22 = r11 if auto_only else (W(sample2, sample2, m2, m2) if do_auto else None)
        return r11, r12, r22

    if normalize_by == "number_counts":
        R11 = plain(sample1, sample1) if do_auto else None
        R12 = R11 if auto_only else (plain(sample1, sample2) if do_cross else None)
        R22 = R11 if auto_only else (plain(sample2, sample2) if do_auto else None)
    elif iterations > 1:
        rows = np.zeros((3, iterations, len(rbins) - 1))
        for it in range(iterations):
            rows[0, it], rows[1, it], rows[2, it] = shuffled()
        R11, R12, R22 = np.median(rows, axis=1)
    else:
        R11, R12, R22 = shuffled()
    return _finish(auto_only, do_auto, do_cross, lambda: W11 / R11, lambda: W12 / R12, lambda: W22 / R22)
