"""GPU pair counters (halotools/mock_observables/pair_counters/__init__.py:1-18 names)."""
from .npairs_3d import npairs_3d  # noqa: F401
from .npairs_projected import npairs_projected  # noqa: F401
from .npairs_xy_z import npairs_xy_z  # noqa: F401
from .marked_npairs_3d import marked_npairs_3d  # noqa: F401
from .marked_npairs_xy_z import marked_npairs_xy_z  # noqa: F401
from .npairs_per_object_3d import npairs_per_object_3d  # noqa: F401
from .npairs_jackknife_3d import npairs_jackknife_3d  # noqa: F401
from .npairs_jackknife_xy_z import npairs_jackknife_xy_z  # noqa: F401
