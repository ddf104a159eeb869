"""Point-set generators restated from the reference's test helpers.

generate_locus_of_3d_points and generate_3d_regular_mesh restate
halotools/mock_observables/tests/cf_helpers.py:12-74.  The reference draws
inside astropy's NumpyRNGContext(seed), i.e. the legacy global RandomState
seeded with `seed`; np.random.RandomState(seed) reproduces those draws.
"""
import numpy as np

fixed_seed = 43


def generate_locus_of_3d_points(npts, xc=0.1, yc=0.1, zc=0.1, epsilon=0.001, seed=None):
    rng = np.random.RandomState(seed)
    x = rng.uniform(xc - epsilon / 2., xc + epsilon / 2., npts)
    y = rng.uniform(yc - epsilon / 2., yc + epsilon / 2., npts)
    z = rng.uniform(zc - epsilon / 2., zc + epsilon / 2., npts)
    return np.vstack([x, y, z]).T


def generate_3d_regular_mesh(npts, dmin=0, dmax=1):
    x = np.linspace(0, 1, npts + 1)
    y = np.linspace(0, 1, npts + 1)
    z = np.linspace(0, 1, npts + 1)
    delta = np.diff(x)[0] / 2.
    x, y, z = np.array(np.meshgrid(x[:-1], y[:-1], z[:-1]))
    return np.vstack([x.flatten() + delta, y.flatten() + delta, z.flatten() + delta]).T


def random_sample(npts, seed=fixed_seed, ncols=3):
    """`with NumpyRNGContext(seed): np.random.random((npts, 3))`."""
    return np.random.RandomState(seed).random_sample((npts, ncols))


def grid_rbins(npts_per_dim=10):
    """The regular-grid rbins of PC/test_pair_counters/test_npairs_3d.py:180-197."""
    grid_spacing = 1. / npts_per_dim
    r1 = grid_spacing / 100.
    epsilon = 0.0001
    r2 = grid_spacing + epsilon
    r3 = grid_spacing * np.sqrt(2) + epsilon
    r4 = grid_spacing * np.sqrt(3) + epsilon
    return np.array([r1, r2, r3, r4])
