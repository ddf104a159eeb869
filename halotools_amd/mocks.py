"""Synthetic point sets and HOD weights for the benchmark configurations.

* ``uniform_box``            -- uniform points in a periodic cube.
* ``nfw_clustered_box``      -- centres uniform in the cube, each with a
  Poisson number of satellites at NFW-distributed radii along isotropic
  directions, wrapped into the box.  The radial law is the NFW enclosed-mass
  fraction P(<x) = g(c x) / g(c), g(y) = ln(1 + y) - y / (1 + y), the profile
  the reference samples in
  halotools/empirical_models/phase_space_models/analytic_models/satellites/nfw/
  kernels/mc_generate_nfw_radial_positions.py:113-128; here it is inverted
  exactly by bisection instead of through a spline table.
* ``zheng07_ncen``           -- the Zheng07 central occupation
  <Ncen> = 0.5 (1 + erf((log10 M - logMmin) / sigma_logM)), as in
  halotools/empirical_models/occupation_models/zheng07_components.py:149-153,
  written in torch so it is differentiable in (logMmin, sigma_logM).
"""
import numpy as np


def uniform_box(n, L, rng):
    return rng.uniform(0, L, (int(n), 3))


def _g(y):
    return np.log1p(y) - y / (1.0 + y)


def nfw_radii(u, conc, rmax, iters=60):
    """Radii r in [0, rmax] with P(< r) = g(c r / rmax) / g(c), by bisection on x = r / rmax."""
    target = u * _g(conc)
    lo = np.zeros_like(u)
    hi = np.ones_like(u)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        below = _g(conc * mid) < target
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return 0.5 * (lo + hi) * rmax


def nfw_clustered_box(ncentres, L, rng, mean_sats=4.0, conc=5.0, rmax=0.5):
    centres = rng.uniform(0, L, (int(ncentres), 3))
    nsat = rng.poisson(mean_sats, int(ncentres))
    host = np.repeat(np.arange(int(ncentres)), nsat)
    r = nfw_radii(rng.uniform(0, 1, len(host)), conc, rmax)
    v = rng.normal(size=(len(host), 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    sats = centres[host] + r[:, None] * v
    pts = np.concatenate([centres, sats])
    # np.mod maps tiny negatives to exactly L; those points sit on the box edge,
    # which the counters accept (the reference clamps x == L into the last cell)
    return np.mod(pts, L)


def zheng07_ncen(log_mass, logMmin, sigma_logM):
    """Zheng07 mean central occupation (torch tensors; differentiable)."""
    import torch
    return 0.5 * (1.0 + torch.erf((log_mass - logMmin) / sigma_logM))


def config_rbins(cfg):
    """Radial edges of the benchmark configurations."""
    if cfg == 1:
        return np.logspace(-1, np.log10(20), 15)
    if cfg == 3:
        return np.logspace(-1, np.log10(30), 16)
    return np.logspace(-1, np.log10(50), 20)


def config_points(cfg, rng=None):
    """Points of configuration 1-4 (seeded with default_rng([43, cfg]))."""
    rng = rng if rng is not None else np.random.default_rng([43, cfg])
    if cfg == 1:
        return uniform_box(20000, 250.0, rng), uniform_box(20000, 250.0, rng), 250.0
    if cfg == 2:
        p = nfw_clustered_box(200000, 1000.0, rng)
        return p, p, 1000.0
    if cfg == 3:
        p = uniform_box(2000000, 1000.0, rng)
        return p, p, 1000.0
    if cfg == 4:
        p = uniform_box(1000000, 1000.0, rng)
        return p, p, 1000.0
    raise ValueError("configurations 1-4 only")


def many_mocks(rng=None, nmocks=64, nmin=5e4, nmax=2e6, L=1000.0):
    """The batched model-fitting workload, seeded with default_rng([43, 5]).

    `nmocks` realisations of N_m points, N_m log-uniform in [nmin, nmax],
    uniform in a periodic cube of side L; halo masses log10 M ~ U[11, 15] per
    point; one Zheng07 logMmin ~ U[11.5, 12.5] per mock; loss coefficients
    c_k ~ N(0, 1), one per edge of ``config_rbins(5)``.  Returns
    (sizes, points list, log-mass list, logMmin array, c array, L).
    """
    rng = rng if rng is not None else np.random.default_rng([43, 5])
    sizes = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nmocks))).astype(int)
    pts = [rng.uniform(0, L, (n, 3)) for n in sizes]
    logm = [rng.uniform(11, 15, n) for n in sizes]
    logMmin = rng.uniform(11.5, 12.5, nmocks)
    c = rng.normal(size=len(config_rbins(5)))
    return sizes, pts, logm, logMmin, c, L
