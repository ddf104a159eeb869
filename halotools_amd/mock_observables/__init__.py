"""Pair counters with the interface of ``halotools.mock_observables``
(halotools/mock_observables/__init__.py:14-15 re-exports the same names)."""
from .pair_counters import (marked_npairs_3d, marked_npairs_xy_z, npairs_3d, npairs_jackknife_3d,  # noqa: F401
                            npairs_jackknife_xy_z, npairs_per_object_3d, npairs_projected, npairs_xy_z)

from .two_point_clustering import marked_tpcf, rp_pi_tpcf, tpcf, wp  # noqa: F401

__all__ = ("npairs_3d", "npairs_projected", "npairs_xy_z", "marked_npairs_3d", "marked_npairs_xy_z",
           "npairs_per_object_3d", "npairs_jackknife_3d", "npairs_jackknife_xy_z", "tpcf", "rp_pi_tpcf", "wp",
           "marked_tpcf")
