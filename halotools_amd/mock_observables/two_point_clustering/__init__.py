"""Two-point functions on the GPU pair counters
(halotools/mock_observables/two_point_clustering/__init__.py names)."""
from .marked_tpcf import marked_tpcf  # noqa: F401
from .rp_pi_tpcf import rp_pi_tpcf  # noqa: F401
from .tpcf import tpcf  # noqa: F401
from .wp import wp  # noqa: F401
