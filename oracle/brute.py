"""Brute-force pair counters in NumPy.

TEST INFRASTRUCTURE ONLY (used by tests/ to pin the oracle).

These restate the reference's own pure-NumPy cross-check helpers in
PC/pairs.py (PC/ = halotools/mock_observables/pair_counters/):

* ``npairs``      -- PC/pairs.py:17-84   (3-D, sqrt(distance) <= r)
* ``xy_z_npairs`` -- PC/pairs.py:87-161  (rp <= rp_bins[i] and |pi| <= pi_bins[j])
* ``wnpairs``     -- PC/pairs.py:164-245 (product-weighted 3-D counts)
* ``distance``    -- PC/pairs.py:445-482 (minimum image via min(|d|, L-|d|))

The method is deliberately different from the engines: it forms every pair,
takes the minimum image with ``min(|dx|, L - |dx|)`` and compares square roots
against the bin edges, so agreement with the engines is evidence, not an echo.
"""
import math

import numpy as np


def _period(period, k):
    if period is None:
        return np.array([np.inf] * k)
    period = np.asarray(period).astype("float64")
    if np.shape(period) == ():
        return np.array([period] * k)
    return period


def _min_image(x1, x2, period):
    d = np.fabs(x1 - x2)
    return np.minimum(d, period - d)


def distance(x1, x2, period=None):
    """PC/pairs.py:445-482."""
    x1 = np.atleast_2d(x1)
    x2 = np.atleast_2d(x2)
    period = _period(period, np.shape(x1)[-1])
    m = _min_image(x1, x2, period)
    return np.sqrt(np.sum(m * m, axis=len(np.shape(m)) - 1))


def npairs(sample1, sample2, rbins, period=None):
    """PC/pairs.py:17-84: number of pairs with separation <= rbins[i]."""
    sample1 = np.atleast_2d(sample1)
    sample2 = np.atleast_2d(sample2)
    rbins = np.atleast_1d(rbins)
    period = _period(period, sample1.shape[-1])
    dd = np.concatenate([distance(sample1[i], sample2, period) for i in range(len(sample1))])
    dd.sort()
    return np.searchsorted(dd, rbins, side="right").astype(np.int64)


def xy_z_npairs(sample1, sample2, rp_bins, pi_bins, period=None):
    """PC/pairs.py:87-161: counts[i, j] of pairs with rp <= rp_bins[i] and pi <= pi_bins[j]."""
    sample1 = np.atleast_2d(sample1)
    sample2 = np.atleast_2d(sample2)
    rp_bins = np.atleast_1d(rp_bins)
    pi_bins = np.atleast_1d(pi_bins)
    period = _period(period, sample1.shape[-1])
    n = np.zeros((rp_bins.size, pi_bins.size), dtype=np.int64)
    for i in range(len(sample1)):
        m = _min_image(sample1[i], sample2, period)
        rp = np.sqrt(np.sum(m[:, :-1] ** 2, axis=1))
        pi = np.sqrt(m[:, -1] ** 2)
        n += ((rp[:, None, None] <= rp_bins[None, :, None]) &
              (pi[:, None, None] <= pi_bins[None, None, :])).sum(axis=0)
    return n


def wnpairs(sample1, sample2, r, period=None, weights1=None, weights2=None):
    """PC/pairs.py:164-245: sum of w1*w2 over pairs with separation <= r[i].

    Accumulated with math.fsum so the brute force is correctly rounded and can
    pin the engines at a relative tolerance far below the reference's 1e-5.
    """
    sample1 = np.atleast_2d(sample1)
    sample2 = np.atleast_2d(sample2)
    r = np.atleast_1d(r)
    period = _period(period, sample1.shape[-1])
    w1 = np.ones(len(sample1)) if weights1 is None else np.asarray(weights1, dtype=np.float64)
    w2 = np.ones(len(sample2)) if weights2 is None else np.asarray(weights2, dtype=np.float64)
    terms = [[] for _ in range(r.size)]
    for i in range(len(sample1)):
        d = distance(sample1[i], sample2, period)
        for k in range(r.size):
            terms[k].extend((w1[i] * w2[d <= r[k]]).tolist())
    return np.array([math.fsum(t) for t in terms])


def wnpairs_generic(sample1, sample2, r, weights1, weights2, wfunc, period=None):
    """Brute-force marked counts for an arbitrary marking function ``wfunc(w1_row, w2_row)``."""
    sample1 = np.atleast_2d(sample1)
    sample2 = np.atleast_2d(sample2)
    r = np.atleast_1d(r)
    period = _period(period, sample1.shape[-1])
    terms = [[] for _ in range(r.size)]
    for i in range(len(sample1)):
        d = distance(sample1[i], sample2, period)
        for j in np.nonzero(d <= r.max())[0]:
            w = wfunc(weights1[i], weights2[j])
            for k in range(r.size):
                if d[j] <= r[k]:
                    terms[k].append(w)
    return np.array([math.fsum(t) for t in terms])


def npairs_per_object(sample1, sample2, rbins, period=None):
    """Per-point counts, the reference's brute-force test helper
    (PC/test_pair_counters/pure_python_npairs_per_object_3d.py:11-47) restated
    with the minimum image and square roots: counts[i, k] = number of sample-2
    points within rbins[k] of sample-1 point i."""
    sample1 = np.atleast_2d(sample1)
    sample2 = np.atleast_2d(sample2)
    rbins = np.atleast_1d(rbins).astype("float64")
    period = _period(period, 3)
    out = np.zeros((len(sample1), len(rbins)), dtype=np.int64)
    for i in range(len(sample1)):
        m = _min_image(sample1[i], sample2, period)
        d = np.sqrt(np.sum(m * m, axis=1))
        out[i] = np.sum(d[:, None] <= rbins[None, :], axis=0)
    return out


def xy_z_wnpairs_generic(sample1, sample2, rp_bins, pi_bins, weights1, weights2, wfunc, period=None):
    """Brute-force marked (rp, pi) sums: PC/pairs.py:248-352 (xy_z_wnpairs, which
    sums w1*w2 over pairs with rp <= rp_bins[i] and pi <= pi_bins[j]) with an
    arbitrary marking function ``wfunc(w1_row, w2_row)``; math.fsum per cell."""
    sample1 = np.atleast_2d(sample1)
    sample2 = np.atleast_2d(sample2)
    rp_bins = np.atleast_1d(rp_bins)
    pi_bins = np.atleast_1d(pi_bins)
    period = _period(period, sample1.shape[-1])
    terms = [[[] for _ in range(pi_bins.size)] for _ in range(rp_bins.size)]
    for i in range(len(sample1)):
        m = _min_image(sample1[i], sample2, period)
        rp = np.sqrt(np.sum(m[:, :-1] ** 2, axis=1))
        pi = np.sqrt(m[:, -1] ** 2)
        for j in np.nonzero((rp <= rp_bins.max()) & (pi <= pi_bins.max()))[0]:
            w = wfunc(weights1[i], weights2[j])
            for a in range(rp_bins.size):
                if rp[j] <= rp_bins[a]:
                    for b in range(pi_bins.size):
                        if pi[j] <= pi_bins[b]:
                            terms[a][b].append(w)
    return np.array([[math.fsum(t) for t in row] for row in terms])
