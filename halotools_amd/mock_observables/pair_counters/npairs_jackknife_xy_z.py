"""``npairs_jackknife_xy_z`` at the reference's module path (PC/npairs_jackknife_xy_z.py);
implemented with ``npairs_jackknife_3d`` in npairs_jackknife_3d.py."""
from .npairs_jackknife_3d import npairs_jackknife_xy_z  # noqa: F401

__all__ = ("npairs_jackknife_xy_z",)
