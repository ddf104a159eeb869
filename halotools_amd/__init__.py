"""halotools_amd -- MI355X-native pair-counting engine for halotools.mock_observables.

Drop-in replacements for ``npairs_3d``, ``npairs_projected``, ``npairs_xy_z`` and
``marked_npairs_3d`` (import them from ``halotools_amd.mock_observables``),
a differentiable product-weighted marked count (``halotools_amd.autodiff``) and a
batched many-mock path (``halotools_amd.batch``).  The counting runs in
hand-written HIP kernels for gfx950 behind the C ABI in include/halopair.h.
"""
__version__ = "0.1.0"
