"""ctypes binding of libhalopair.so (the C ABI in include/halopair.h).

The shared library is built in-tree by ``halotools_amd/build_ext.py`` (or
``__graft_entry__.build()``).  There is no CPU fallback: if the library or a
ROCm GPU is missing, every counting call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HALOPAIR_LIB selects another build of the same ABI (A/B comparisons of kernel variants)
LIB_PATH = os.environ.get("HALOPAIR_LIB") or os.path.join(_HERE, "_lib", "libhalopair.so")

HP_NPAIRS_3D = 0
HP_NPAIRS_PROJECTED = 1
HP_NPAIRS_XY_Z = 2
HP_MARKED_NPAIRS_3D = 3
HP_NPAIRS_PER_OBJECT_3D = 4
HP_MARKED_NPAIRS_XY_Z = 5
HP_NPAIRS_JACKKNIFE_3D = 6
HP_NPAIRS_JACKKNIFE_XY_Z = 7

HP_SAME_SAMPLES = 1
HP_ROWS1 = 2
HP_ROWS2 = 4
HP_NO_SUMS = 8
HP_ALSO_COUNTS = 16
HP_TANGENT = 32

EXPORTED_SYMBOLS = (
    "hp_last_error", "hp_abi_version", "hp_workspace_size", "hp_npairs_3d",
    "hp_npairs_projected", "hp_npairs_xy_z", "hp_marked_npairs_3d", "hp_npairs_per_object_3d",
    "hp_marked_npairs_xy_z", "hp_npairs_jackknife_3d", "hp_npairs_jackknife_xy_z", "hp_marked_backward",
    "hp_segment_sum_scratch", "hp_segment_sum", "hp_debug_bin_lookup", "hp_debug_hilbert",
    "hp_npairs_3d_noalloc", "hp_zheng07_tangent",
)


class HpMock(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int64),
        ("n2", ctypes.c_int64),
        ("period", ctypes.c_double * 3),
        ("pbc", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("cell", ctypes.c_double * 3),
    ]


class HpRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("nmocks", ctypes.c_int32),
        ("wfunc", ctypes.c_int32),
        ("mocks", ctypes.POINTER(HpMock)),
        ("x1", ctypes.c_void_p), ("y1", ctypes.c_void_p), ("z1", ctypes.c_void_p),
        ("x2", ctypes.c_void_p), ("y2", ctypes.c_void_p), ("z2", ctypes.c_void_p),
        ("w1", ctypes.c_void_p), ("w2", ctypes.c_void_p),
        ("nw", ctypes.c_int32),
        ("nr", ctypes.c_int32),
        ("np", ctypes.c_int32),
        ("tile_waves", ctypes.c_int32),
        ("r2", ctypes.POINTER(ctypes.c_double)),
        ("p2", ctypes.POINTER(ctypes.c_double)),
        ("search", ctypes.c_double * 3),
        ("counts", ctypes.c_void_p),
        ("sums", ctypes.c_void_p),
        ("rows1", ctypes.c_void_p),
        ("rows2", ctypes.c_void_p),
        ("evals", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_size_t),
        ("stream", ctypes.c_void_p),
        ("pair_events", ctypes.c_void_p * 2),
        ("build_events", ctypes.c_void_p * 2),
        ("tangents", ctypes.c_void_p),
        ("jtags1", ctypes.c_void_p),
        ("jtags2", ctypes.c_void_p),
        ("n_samples", ctypes.c_int32),
        ("stride1", ctypes.c_int32),
        ("stride2", ctypes.c_int32),
        ("reserved2", ctypes.c_int32),
    ]


_LIB = None


class EngineError(RuntimeError):
    """Raised when libhalopair.so reports an error."""


def library():
    """Load libhalopair.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libhalopair.so is not built (%s); run `python halotools_amd/build_ext.py` "
            "or __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.hp_last_error.restype = ctypes.c_char_p
    lib.hp_abi_version.restype = ctypes.c_int
    lib.hp_workspace_size.argtypes = [ctypes.POINTER(HpRequest)]
    lib.hp_workspace_size.restype = ctypes.c_size_t
    for name in ("hp_npairs_3d", "hp_npairs_projected", "hp_npairs_xy_z", "hp_marked_npairs_3d",
                 "hp_npairs_per_object_3d", "hp_marked_npairs_xy_z", "hp_npairs_jackknife_3d",
                 "hp_npairs_jackknife_xy_z", "hp_npairs_3d_noalloc"):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.POINTER(HpRequest)]
        fn.restype = ctypes.c_int
    lib.hp_marked_backward.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_int64), ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p]
    lib.hp_marked_backward.restype = ctypes.c_int
    lib.hp_segment_sum_scratch.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    lib.hp_segment_sum_scratch.restype = ctypes.c_int64
    lib.hp_segment_sum.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    lib.hp_segment_sum.restype = ctypes.c_int
    lib.hp_zheng07_tangent.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    lib.hp_zheng07_tangent.restype = ctypes.c_int
    lib.hp_debug_bin_lookup.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.hp_debug_bin_lookup.restype = ctypes.c_int
    lib.hp_debug_hilbert.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32]
    lib.hp_debug_hilbert.restype = ctypes.c_uint32
    _LIB = lib
    return lib


def check(rc):
    if rc != 0:
        msg = library().hp_last_error().decode()
        raise EngineError("libhalopair: " + (msg or "error %d" % rc))


def debug_bin_lookup(r2, d):
    """Host evaluation of the kernels' bucketed bin lookup (first bin index of each d)."""
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.zeros(len(d), dtype=np.int32)
    info = np.zeros(4, dtype=np.int32)
    P = ctypes.POINTER
    check(library().hp_debug_bin_lookup(
        r2.ctypes.data_as(P(ctypes.c_double)), len(r2), d.ctypes.data_as(P(ctypes.c_double)),
        len(d), out.ctypes.data_as(P(ctypes.c_int32)), info.ctypes.data_as(P(ctypes.c_int32))))
    return out, dict(shift=int(info[0]), base=int(info[1]), nb=int(info[2]), multi=int(info[3]))
