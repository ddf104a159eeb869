"""CPU oracle: NumPy restatement of the reference mesh + ctypes driver for cpu_ref.c.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke`` and by
the ``cpu_baseline`` leg of bench.py -- never by the product package.

This module restates, function for function, the pieces of the reference that
build the cell-list mesh and feed the Cython engines:

* ``digitized_position``      -- PC/rectangular_mesh.py:19-22
* ``sample1_cell_size``       -- PC/rectangular_mesh.py:25-55
* ``sample2_cell_sizes``      -- PC/rectangular_mesh.py:58-82
* ``RefMesh``                 -- PC/rectangular_mesh.py:85-225 (RectangularMesh)
* ``RefDoubleMesh``           -- PC/rectangular_mesh.py:228-417
* ``enclose_in_box``          -- PC/mesh_helpers.py:17-64
* ``cell1_parallelization``   -- PC/mesh_helpers.py:183-221
* ``npairs_3d`` / ``npairs_projected`` / ``npairs_xy_z`` / ``marked_npairs_3d``
  -- the control flow of PC/npairs_3d.py:111-150, PC/npairs_projected.py:116-157,
  PC/npairs_xy_z.py:122-163, PC/marked_npairs_3d.py:141-183, with the engine loop
  itself in cpu_ref.c.

PC/ abbreviates halotools/mock_observables/pair_counters/.  Argument *validation*
messages are not restated here (they live in the product and are pinned by their
own tests); the oracle assumes valid input.
"""
import ctypes
import os
from math import floor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

default_max_cells_per_dimension_cell1 = 50
default_max_cells_per_dimension_cell2 = 50


class _RefMeshC(ctypes.Structure):
    _fields_ = [
        ("first_cell1", ctypes.c_int64),
        ("last_cell1", ctypes.c_int64),
        ("n1div", ctypes.c_int32 * 3),
        ("n2div", ctypes.c_int32 * 3),
        ("ncover", ctypes.c_int32 * 3),
        ("pbc", ctypes.c_int32),
        ("period", ctypes.c_double * 3),
        ("cell1_indices", ctypes.c_void_p),
        ("cell2_indices", ctypes.c_void_p),
    ]


_LIBS = {}


def build():
    """Compile cpu_ref.c (exact and -Ofast flavours) with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load(fast=False):
    """Load the exact-mode (parity) or -Ofast (baseline) oracle library."""
    name = "libhp_oracle_fast.so" if fast else "libhp_oracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_BUILD, name)
    if not fast and os.environ.get("HP_ORACLE_LIB"):
        path = os.environ["HP_ORACLE_LIB"]  # e.g. the sanitizer build (make -C oracle san)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    D = ctypes.c_double
    I = ctypes.c_int
    lib.ref_npairs_3d.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, I, I, P]
    lib.ref_npairs_projected.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, I, D, I, P]
    lib.ref_npairs_xy_z.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, I, P, I, I, P]
    lib.ref_marked_npairs_3d.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, P, I, I, P, I, I, P]
    lib.ref_marked_npairs_xy_z.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, P, I, I, P, I, P, I, I,
                                           P]
    lib.ref_npairs_per_object_3d.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, I, ctypes.c_int64, I, P]
    lib.ref_npairs_jackknife.argtypes = [ctypes.POINTER(_RefMeshC), P, P, P, P, P, P, P, P, P, P, ctypes.c_int64, P,
                                         I, P, I, I, P]
    lib.ref_candidate_pairs.argtypes = [ctypes.POINTER(_RefMeshC)]
    lib.ref_candidate_pairs.restype = ctypes.c_int64
    lib.ref_set_mutation.argtypes = [I]
    lib.ref_set_compensated.argtypes = [I]
    lib.ref_wfunc.argtypes = [I, P, P]
    lib.ref_wfunc.restype = D
    _LIBS[name] = lib
    return lib


def set_compensated(on):
    """Compensated (Neumaier) summation of the exact oracle's marked sums (tests only)."""
    load().ref_set_compensated(1 if on else 0)


def set_mutation(m):
    """Plant a known mistake in the exact-mode oracle (tests only); 0 clears it."""
    load().ref_set_mutation(int(m))


# --------------------------------------------------------------------------
# PC/rectangular_mesh.py
# --------------------------------------------------------------------------

def digitized_position(p, cell_size, num_divs):
    """PC/rectangular_mesh.py:19-22 -- floor(p // cs), points at/after the end clamp to n-1."""
    ip = np.floor(p // cell_size).astype(int)
    return np.where(ip >= num_divs, num_divs - 1, ip)


def sample1_cell_size(period, search_length, approx_cell_size,
                      max_cells_per_dimension=default_max_cells_per_dimension_cell1):
    """PC/rectangular_mesh.py:25-55."""
    period = float(period)
    if search_length > period / 3.0:
        raise ValueError("Input ``search_length`` cannot exceed period/3")
    ndivs = int(floor(period / float(approx_cell_size)))
    ndivs = max(ndivs, 1)
    ndivs = min(max_cells_per_dimension, ndivs)
    nsearch = int(floor(period / float(search_length)))
    nsearch = max(nsearch, 1)
    ndivs = min(ndivs, nsearch)
    ndivs = max(3, ndivs)
    return period / float(ndivs)


def sample2_cell_sizes(period, sample1_cell_size, approx_cell_size,
                       max_cells_per_dimension=default_max_cells_per_dimension_cell2):
    """PC/rectangular_mesh.py:58-82."""
    period = float(period)
    num_sample1_cells = int(np.round(period / sample1_cell_size))
    ndivs_sample1_cells = int(np.round(sample1_cell_size / float(approx_cell_size)))
    ndivs_sample1_cells = max(1, ndivs_sample1_cells)
    ndivs_sample1_cells = min(max_cells_per_dimension, ndivs_sample1_cells)
    num_sample2_cells = num_sample1_cells * ndivs_sample1_cells
    if num_sample2_cells > max_cells_per_dimension:
        num2_per_num1 = max_cells_per_dimension // num_sample1_cells
        num_sample2_cells = num2_per_num1 * num_sample1_cells
    return period / float(num_sample2_cells)


class RefMesh(object):
    """PC/rectangular_mesh.py:85-225 (RectangularMesh).

    ``argsort`` uses kind='stable'; the reference uses NumPy's default
    (unstable) sort, which only changes the order of points inside a cell, and
    hence only the floating-point summation order of marked sums.
    """

    def __init__(self, x, y, z, xperiod, yperiod, zperiod, cx, cy, cz):
        self.npts = x.shape[0]
        self.xperiod, self.yperiod, self.zperiod = float(xperiod), float(yperiod), float(zperiod)
        self.num_xdivs = max(int(np.round(self.xperiod / cx)), 1)
        self.num_ydivs = max(int(np.round(self.yperiod / cy)), 1)
        self.num_zdivs = max(int(np.round(self.zperiod / cz)), 1)
        self.ncells = self.num_xdivs * self.num_ydivs * self.num_zdivs
        self.xcell_size = self.xperiod / float(self.num_xdivs)
        self.ycell_size = self.yperiod / float(self.num_ydivs)
        self.zcell_size = self.zperiod / float(self.num_zdivs)
        ix = digitized_position(x, self.xcell_size, self.num_xdivs)
        iy = digitized_position(y, self.ycell_size, self.num_ydivs)
        iz = digitized_position(z, self.zcell_size, self.num_zdivs)
        cell_ids = ix * (self.num_ydivs * self.num_zdivs) + iy * self.num_zdivs + iz
        self.idx_sorted = np.ascontiguousarray(np.argsort(cell_ids, kind="stable"))
        cell_id_indices = np.searchsorted(cell_ids, np.arange(self.ncells), sorter=self.idx_sorted)
        self.cell_id_indices = np.ascontiguousarray(np.append(cell_id_indices, self.npts).astype(np.int64))


class RefDoubleMesh(object):
    """PC/rectangular_mesh.py:228-374 (RectangularDoubleMesh)."""

    def __init__(self, x1, y1, z1, x2, y2, z2, a1, a2, search, period, PBCs=True,
                 max1=default_max_cells_per_dimension_cell1,
                 max2=default_max_cells_per_dimension_cell2):
        self.xperiod, self.yperiod, self.zperiod = [float(p) for p in period]
        self.search_xlength, self.search_ylength, self.search_zlength = search
        self._PBCs = PBCs
        for s, p in zip(search, period):
            if not s <= p / 3.0:
                raise ValueError("search length exceeds period/3")
        c1 = [sample1_cell_size(p, s, a, max1) for p, s, a in zip(period, search, a1)]
        self.mesh1 = RefMesh(x1, y1, z1, *period, *c1)
        c2 = [sample2_cell_sizes(p, cs, a, max2) for p, cs, a in
              zip(period, (self.mesh1.xcell_size, self.mesh1.ycell_size, self.mesh1.zcell_size), a2)]
        self.mesh2 = RefMesh(x2, y2, z2, *period, *c2)

    def covering_steps(self):
        """int(np.ceil(search / cell2)) -- PC/cpairs/npairs_3d_engine.pyx:74-79."""
        return (int(np.ceil(self.search_xlength / self.mesh2.xcell_size)),
                int(np.ceil(self.search_ylength / self.mesh2.ycell_size)),
                int(np.ceil(self.search_zlength / self.mesh2.zcell_size)))

    def as_c(self, first=0, last=None):
        m1, m2 = self.mesh1, self.mesh2
        s = _RefMeshC()
        s.first_cell1 = int(first)
        s.last_cell1 = int(m1.ncells if last is None else last)
        s.n1div[:] = [m1.num_xdivs, m1.num_ydivs, m1.num_zdivs]
        s.n2div[:] = [m2.num_xdivs, m2.num_ydivs, m2.num_zdivs]
        s.ncover[:] = list(self.covering_steps())
        s.pbc = int(bool(self._PBCs))
        s.period[:] = [self.xperiod, self.yperiod, self.zperiod]
        s.cell1_indices = m1.cell_id_indices.ctypes.data
        s.cell2_indices = m2.cell_id_indices.ctypes.data
        return s


# --------------------------------------------------------------------------
# PC/mesh_helpers.py
# --------------------------------------------------------------------------

def enclose_in_box(x1, y1, z1, x2, y2, z2, min_size=None):
    """PC/mesh_helpers.py:17-64: one scalar shift (the min over all six arrays)."""
    xyzmin = np.min([np.min(x1), np.min(x2), np.min(y1), np.min(y2), np.min(z1), np.min(z2)])
    xyzmax = np.max([np.max(x1), np.max(x2), np.max(y1), np.max(y2), np.max(z1), np.max(z2)]) - xyzmin
    x1, y1, z1 = x1 - xyzmin, y1 - xyzmin, z1 - xyzmin
    x2, y2, z2 = x2 - xyzmin, y2 - xyzmin, z2 - xyzmin
    Lbox = np.array([xyzmax, xyzmax, xyzmax])
    if min_size is not None:
        min_size = np.atleast_1d(min_size)
        if np.any(Lbox < min_size):
            Lbox[(Lbox < min_size)] = min_size[(Lbox < min_size)]
    return x1, y1, z1, x2, y2, z2, Lbox


def cell1_parallelization(ncells, num_threads):
    """PC/mesh_helpers.py:183-221 (returns the tuples; the C driver does the split itself)."""
    if num_threads == 1:
        return 1, [(0, ncells)]
    elif num_threads > ncells:
        return ncells, [(a, a + 1) for a in np.arange(ncells)]
    parts = [a for a in np.array_split(np.arange(ncells), num_threads) if len(a) > 0]
    return num_threads, [(x[0], x[0] + len(x)) for x in parts]


# --------------------------------------------------------------------------
# drivers
# --------------------------------------------------------------------------

def _cols(sample):
    s = np.asarray(sample)
    return s[:, 0], s[:, 1], s[:, 2]


def _period3(period):
    p = np.atleast_1d(period).astype(float)
    if len(p) == 1:
        p = np.array([p[0]] * 3)
    return p


def _cells3(a, default):
    if a is None:
        return list(default)
    a = np.atleast_1d(a).astype(float)
    if len(a) == 1:
        return [a[0]] * 3
    return list(a)


def _prepare(sample1, sample2, search, period, approx_cell1_size, approx_cell2_size, default_cells):
    x1, y1, z1 = _cols(sample1)
    x2, y2, z2 = _cols(sample2)
    if period is None:
        PBCs = False
        x1, y1, z1, x2, y2, z2, period = enclose_in_box(
            x1, y1, z1, x2, y2, z2, min_size=[3.0 * s for s in search])
    else:
        PBCs = True
        period = _period3(period)
    a1 = _cells3(approx_cell1_size, default_cells)
    a2 = _cells3(approx_cell2_size, default_cells)
    mesh = RefDoubleMesh(x1, y1, z1, x2, y2, z2, a1, a2, search, period, PBCs)
    i1, i2 = mesh.mesh1.idx_sorted, mesh.mesh2.idx_sorted
    c = lambda a, idx: np.ascontiguousarray(np.asarray(a, dtype=np.float64)[idx])
    sorted_cols = (c(x1, i1), c(y1, i1), c(z1, i1), c(x2, i2), c(y2, i2), c(z2, i2))
    return mesh, sorted_cols


def _ptrs(cols):
    return [a.ctypes.data for a in cols]


def npairs_3d(sample1, sample2, rbins, period=None, num_threads=1,
              approx_cell1_size=None, approx_cell2_size=None, fast=False):
    """Oracle for PC/npairs_3d.py:20-150 (engine loop in cpu_ref.c)."""
    rbins = np.atleast_1d(rbins).astype("f8")
    rmax = np.max(rbins)
    mesh, cols = _prepare(sample1, sample2, (rmax, rmax, rmax), period,
                          approx_cell1_size, approx_cell2_size, (rmax, rmax, rmax))
    counts = np.zeros(len(rbins), dtype=np.int64)
    mc = mesh.as_c()
    rc = load(fast).ref_npairs_3d(ctypes.byref(mc), *_ptrs(cols), rbins.ctypes.data,
                                  len(rbins), int(num_threads), counts.ctypes.data)
    assert rc == 0
    return counts


def npairs_3d_phases(sample1, sample2, rbins, period=None, num_threads=1, fast=False):
    """npairs_3d with its two phases timed apart (for the CPU baseline):
    returns (counts, seconds of the mesh build and argument handling,
    seconds of the engine loop)."""
    import time
    t0 = time.perf_counter()
    rbins = np.atleast_1d(rbins).astype("f8")
    rmax = np.max(rbins)
    mesh, cols = _prepare(sample1, sample2, (rmax, rmax, rmax), period, None, None, (rmax, rmax, rmax))
    counts = np.zeros(len(rbins), dtype=np.int64)
    mc = mesh.as_c()
    ptrs = _ptrs(cols)
    lib = load(fast)
    t1 = time.perf_counter()
    rc = lib.ref_npairs_3d(ctypes.byref(mc), *ptrs, rbins.ctypes.data, len(rbins), int(num_threads),
                           counts.ctypes.data)
    t2 = time.perf_counter()
    assert rc == 0
    return counts, t1 - t0, t2 - t1


def npairs_per_object_3d(sample1, sample2, rbins, period=None, num_threads=1,
                         approx_cell1_size=None, approx_cell2_size=None, fast=False):
    """Oracle for PC/npairs_per_object_3d.py:20-137: (Npts1, K) counts per
    sample-1 point, unsorted back to input order like the engine
    (PC/cpairs/npairs_per_object_3d_engine.pyx:202-205)."""
    rbins = np.atleast_1d(rbins).astype("f8")
    rmax = np.max(rbins)
    mesh, cols = _prepare(sample1, sample2, (rmax, rmax, rmax), period,
                          approx_cell1_size, approx_cell2_size, (rmax, rmax, rmax))
    n1 = len(cols[0])
    sorted_counts = np.zeros((n1, len(rbins)), dtype=np.int64)
    mc = mesh.as_c()
    rc = load(fast).ref_npairs_per_object_3d(ctypes.byref(mc), *_ptrs(cols), rbins.ctypes.data, len(rbins),
                                             n1, int(num_threads), sorted_counts.ctypes.data)
    assert rc == 0
    counts = np.empty_like(sorted_counts)
    counts[mesh.mesh1.idx_sorted] = sorted_counts
    return counts


def npairs_projected(sample1, sample2, rp_bins, pi_max, period=None, num_threads=1,
                     approx_cell1_size=None, approx_cell2_size=None, fast=False):
    """Oracle for PC/npairs_projected.py:20-157."""
    rp_bins = np.atleast_1d(rp_bins).astype("f8")
    rp_max = np.max(rp_bins)
    pi_max = float(pi_max)
    mesh, cols = _prepare(sample1, sample2, (rp_max, rp_max, pi_max), period,
                          approx_cell1_size, approx_cell2_size, (rp_max, rp_max, rp_max))
    counts = np.zeros(len(rp_bins), dtype=np.int64)
    mc = mesh.as_c()
    rc = load(fast).ref_npairs_projected(ctypes.byref(mc), *_ptrs(cols), rp_bins.ctypes.data,
                                         len(rp_bins), pi_max, int(num_threads), counts.ctypes.data)
    assert rc == 0
    return counts


def npairs_xy_z(sample1, sample2, rp_bins, pi_bins, period=None, num_threads=1,
                approx_cell1_size=None, approx_cell2_size=None, fast=False):
    """Oracle for PC/npairs_xy_z.py:20-163."""
    rp_bins = np.atleast_1d(rp_bins).astype("f8")
    pi_bins = np.atleast_1d(pi_bins).astype("f8")
    rp_max, pi_max = np.max(rp_bins), np.max(pi_bins)
    mesh, cols = _prepare(sample1, sample2, (rp_max, rp_max, pi_max), period,
                          approx_cell1_size, approx_cell2_size, (rp_max, rp_max, pi_max))
    counts = np.zeros((len(rp_bins), len(pi_bins)), dtype=np.int64)
    mc = mesh.as_c()
    rc = load(fast).ref_npairs_xy_z(ctypes.byref(mc), *_ptrs(cols), rp_bins.ctypes.data, len(rp_bins),
                                    pi_bins.ctypes.data, len(pi_bins), int(num_threads), counts.ctypes.data)
    assert rc == 0
    return counts


def marked_npairs_3d(sample1, sample2, rbins, period=None, weights1=None, weights2=None,
                     weight_func_id=0, num_threads=1, approx_cell1_size=None,
                     approx_cell2_size=None, fast=False, nw=None):
    """Oracle for PC/marked_npairs_3d.py:25-183 (weights given as (N, nw) or (N,))."""
    rbins = np.atleast_1d(rbins).astype("f8")
    rmax = np.max(rbins)
    n1, n2 = len(sample1), len(sample2)
    if nw is None:
        nw = 1 if weights1 is None or np.ndim(weights1) == 1 else np.shape(weights1)[1]
    w1 = np.ones((n1, nw)) if weights1 is None else np.asarray(weights1, dtype=np.float64).reshape(n1, -1)
    w2 = np.ones((n2, nw)) if weights2 is None else np.asarray(weights2, dtype=np.float64).reshape(n2, -1)
    mesh, cols = _prepare(sample1, sample2, (rmax, rmax, rmax), period,
                          approx_cell1_size, approx_cell2_size, (rmax, rmax, rmax))
    w1s = np.ascontiguousarray(w1[mesh.mesh1.idx_sorted, :])
    w2s = np.ascontiguousarray(w2[mesh.mesh2.idx_sorted, :])
    out = np.zeros(len(rbins), dtype=np.float64)
    mc = mesh.as_c()
    rc = load(fast).ref_marked_npairs_3d(ctypes.byref(mc), *_ptrs(cols), w1s.ctypes.data, w2s.ctypes.data,
                                         w1s.shape[1], int(weight_func_id), rbins.ctypes.data, len(rbins),
                                         int(num_threads), out.ctypes.data)
    assert rc == 0
    return out


def marked_npairs_xy_z(sample1, sample2, rp_bins, pi_bins, period=None, weights1=None, weights2=None,
                       weight_func_id=0, num_threads=1, approx_cell1_size=None, approx_cell2_size=None,
                       fast=False, nw=None):
    """Oracle for PC/marked_npairs_xy_z.py:23-194: (len(rp_bins), len(pi_bins)) sums of
    wfunc(w1_i, w2_j) (engine loop PC/marked_cpairs/marked_npairs_xy_z_engine.pyx:173-226)."""
    rp_bins = np.atleast_1d(rp_bins).astype("f8")
    pi_bins = np.atleast_1d(pi_bins).astype("f8")
    rp_max, pi_max = np.max(rp_bins), np.max(pi_bins)
    n1, n2 = len(sample1), len(sample2)
    if nw is None:
        nw = 1 if weights1 is None or np.ndim(weights1) == 1 else np.shape(weights1)[1]
    w1 = np.ones((n1, nw)) if weights1 is None else np.asarray(weights1, dtype=np.float64).reshape(n1, -1)
    w2 = np.ones((n2, nw)) if weights2 is None else np.asarray(weights2, dtype=np.float64).reshape(n2, -1)
    mesh, cols = _prepare(sample1, sample2, (rp_max, rp_max, pi_max), period,
                          approx_cell1_size, approx_cell2_size, (rp_max, rp_max, pi_max))
    w1s = np.ascontiguousarray(w1[mesh.mesh1.idx_sorted, :])
    w2s = np.ascontiguousarray(w2[mesh.mesh2.idx_sorted, :])
    out = np.zeros((len(rp_bins), len(pi_bins)), dtype=np.float64)
    mc = mesh.as_c()
    rc = load(fast).ref_marked_npairs_xy_z(ctypes.byref(mc), *_ptrs(cols), w1s.ctypes.data, w2s.ctypes.data,
                                           w1s.shape[1], int(weight_func_id), rp_bins.ctypes.data, len(rp_bins),
                                           pi_bins.ctypes.data, len(pi_bins), int(num_threads), out.ctypes.data)
    assert rc == 0
    return out


def npairs_jackknife(sample1, sample2, rp_bins, jtags1, jtags2, N_samples, pi_bins=None, period=None,
                     weights1=None, weights2=None, num_threads=1, approx_cell1_size=None, approx_cell2_size=None,
                     fast=False):
    """Oracle for PC/npairs_jackknife_3d.py:24-95 (pi_bins None) and
    PC/npairs_jackknife_xy_z.py:23-199: (N_samples+1, K) or (N_samples+1, K, G)
    float64 (engine loops PC/cpairs/npairs_jackknife_3d_engine.pyx:98-236,
    PC/cpairs/npairs_jackknife_xy_z_engine.pyx:98-246)."""
    rp_bins = np.atleast_1d(rp_bins).astype("f8")
    rmax = np.max(rp_bins)
    zmax = rmax if pi_bins is None else np.max(pi_bins)
    if pi_bins is not None:
        pi_bins = np.atleast_1d(pi_bins).astype("f8")
    n1, n2 = len(sample1), len(sample2)
    w1 = np.ones(n1) if weights1 is None else np.asarray(weights1, dtype=np.float64)
    w2 = np.ones(n2) if weights2 is None else np.asarray(weights2, dtype=np.float64)
    jt1 = np.asarray(jtags1).astype(np.int64)
    jt2 = np.asarray(jtags2).astype(np.int64)
    mesh, cols = _prepare(sample1, sample2, (rmax, rmax, zmax), period,
                          approx_cell1_size, approx_cell2_size, (rmax, rmax, zmax))
    w1s = np.ascontiguousarray(w1[mesh.mesh1.idx_sorted])
    w2s = np.ascontiguousarray(w2[mesh.mesh2.idx_sorted])
    j1s = np.ascontiguousarray(jt1[mesh.mesh1.idx_sorted])
    j2s = np.ascontiguousarray(jt2[mesh.mesh2.idx_sorted])
    G = 1 if pi_bins is None else len(pi_bins)
    out = np.zeros((N_samples + 1, len(rp_bins), G), dtype=np.float64)
    mc = mesh.as_c()
    rc = load(fast).ref_npairs_jackknife(ctypes.byref(mc), *_ptrs(cols), w1s.ctypes.data, w2s.ctypes.data,
                                     j1s.ctypes.data, j2s.ctypes.data, int(N_samples), rp_bins.ctypes.data,
                                     len(rp_bins), None if pi_bins is None else pi_bins.ctypes.data, G,
                                     int(num_threads), out.ctypes.data)
    assert rc == 0
    return out[:, :, 0] if pi_bins is None else out


def candidate_pairs_3d(sample1, sample2, rmax, period, approx_cell1_size=None, approx_cell2_size=None):
    """Number of candidate pair evaluations the reference engine performs (its own mesh)."""
    mesh, _ = _prepare(sample1, sample2, (rmax, rmax, rmax), period,
                       approx_cell1_size, approx_cell2_size, (rmax, rmax, rmax))
    mc = mesh.as_c()
    return int(load().ref_candidate_pairs(ctypes.byref(mc)))


def wfunc(id_, w1, w2):
    """Evaluate one restated marking function on two weight rows."""
    a = np.ascontiguousarray(w1, dtype=np.float64)
    b = np.ascontiguousarray(w2, dtype=np.float64)
    return load().ref_wfunc(int(id_), a.ctypes.data, b.ctypes.data)
