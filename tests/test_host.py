"""CPU-side tests of the product's host logic: the C ABI library loads and
exports every symbol of include/halopair.h, the bucketed bin lookup that the
kernels use is exact, and argument validation raises the reference's errors
(PC/test_pair_counters/test_npairs_3d.py:270-304,
PC/test_pair_counters/test_marked_npairs_3d.py:181-232,
PC/test_pair_counters/test_mesh_helpers.py:10-79)."""
import ctypes
import os
import re

import numpy as np
import pytest

from halotools_amd import _engine as E
from halotools_amd.custom_exceptions import HalotoolsError
from halotools_amd.mock_observables import marked_npairs_3d, marked_npairs_xy_z, npairs_3d, npairs_projected, npairs_xy_z
from halotools_amd.mock_observables.pair_counters import _process as P
from halotools_amd.mock_observables.pair_counters.marked_npairs_3d import _func_signature_int_from_wfunc
from tests.helpers import fixed_seed, generate_locus_of_3d_points

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "halopair.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(hp_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = E.library()
    declared = _header_functions()
    assert "hp_npairs_3d" in declared and "hp_marked_backward" in declared
    for name in declared:
        assert hasattr(lib, name), name
        ctypes.cast(getattr(lib, name), ctypes.c_void_p)
    assert lib.hp_abi_version() == 2


def test_declared_symbols_match_binding_list():
    assert set(_header_functions()) <= set(E.EXPORTED_SYMBOLS)


def test_abi_struct_sizes():
    # hp_mock and hp_request must match the C layout of include/halopair.h
    assert ctypes.sizeof(E.HpMock) == 8 + 8 + 24 + 4 + 4 + 24
    off = E.HpRequest.workspace.offset
    assert off % 8 == 0 and ctypes.sizeof(E.HpRequest) == E.HpRequest.reserved2.offset + 4


def test_abi_layout_matches_c_compiler(tmp_path):
    # the ctypes mirror against what a C compiler makes of include/halopair.h
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "halopair.h"', "int main(void) {"]
    for cname, cls in (("hp_mock", E.HpMock), ("hp_request", E.HpRequest)):
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for field in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, field[0], cname, field[0]))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.splitlines())
    for cname, cls in (("hp_mock", E.HpMock), ("hp_request", E.HpRequest)):
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for field in cls._fields_:
            assert int(got["%s.%s" % (cname, field[0])]) == getattr(cls, field[0]).offset, (cname, field[0])


@pytest.mark.parametrize("edges", [
    np.logspace(-1, np.log10(50), 20),
    np.logspace(-1, np.log10(30), 16),
    np.arange(0, 41, 1.0),
    np.array([0.001, 0.1, 0.2, 0.3]),
    np.linspace(0.001, 0.3, 100),
    np.array([0.0, 0.1, 0.2, 0.3]),
    np.array([1.0, 1.0 + 1e-12, 2.0]),          # two edges inside one bucket
    np.array([0.0, 0.0]),
    np.logspace(-8, 8, 60),
])
def test_bin_lookup_matches_searchsorted(edges):
    r2 = edges * edges
    rng = np.random.default_rng(0)
    d = np.concatenate([r2, np.nextafter(r2, -1), np.nextafter(r2, np.inf), [0.0, 5e-324, 1e-300, 1e300, np.inf],
                        rng.uniform(0, r2.max() * 1.5 + 1, 20000), np.exp(rng.uniform(-40, 20, 20000))])
    d = d[d >= 0]
    out, info = E.debug_bin_lookup(r2, d)
    e = np.minimum.accumulate(r2[::-1])[::-1]
    ref = np.searchsorted(e, d, side="left")
    assert np.array_equal(out, ref), info
    out_nan, _ = E.debug_bin_lookup(r2, np.array([np.nan]))
    assert out_nan[0] == len(r2)


def test_bin_lookup_two_element_descending():
    # rbins = [b, a] with b > a: the reference counts a pair in both bins iff dsq <= a**2
    r2 = np.array([0.04, 0.01])
    out, _ = E.debug_bin_lookup(r2, np.array([0.005, 0.01, 0.02, 0.04, 0.05]))
    assert list(out) == [0, 0, 2, 2, 2]


# ------------------------------------------------------------- validation (no GPU needed)
def _loci():
    d1 = generate_locus_of_3d_points(100, xc=0.1, yc=0.1, zc=0.1, seed=fixed_seed)
    d2 = generate_locus_of_3d_points(100, xc=0.1, yc=0.1, zc=0.2, seed=fixed_seed)
    return d1, d2


def test_sensible_num_threads():
    d1, d2 = _loci()
    with pytest.raises(ValueError) as err:
        npairs_3d(d1, d2, np.array((0.05, 0.15, 0.3)), period=1, num_threads="Cuba Gooding Jr.")
    assert "Input ``num_threads`` argument must be an integer or the string 'max'" in err.value.args[0]


def test_sensible_rbins():
    d1, d2 = _loci()
    with pytest.raises(ValueError) as err:
        npairs_3d(d1, d2, 0.1, period=1)
    assert "Input ``rbins`` must be a monotonically increasing 1D array with at least two entries" in err.value.args[0]
    with pytest.raises(ValueError):
        npairs_3d(d1, d2, np.array([0.1, 0.3, 0.2]), period=1)


def test_sensible_period():
    d1, d2 = _loci()
    with pytest.raises(ValueError) as err:
        npairs_3d(d1, d2, np.array((0.05, 0.15, 0.3)), period=np.inf)
    assert "Input ``period`` must be a bounded positive number in all dimensions" in err.value.args[0]


def test_search_length_limit():
    d1, d2 = _loci()
    with pytest.raises(ValueError) as err:
        npairs_3d(d1, d2, np.array((0.05, 0.15, 0.4)), period=1)
    assert "The maximum length over which you search for pairs of points" in err.value.args[0]


def test_rp_pi_bins_messages():
    d1, d2 = _loci()
    with pytest.raises(ValueError) as err:
        npairs_xy_z(d1, d2, np.array([0.1]), np.array([0.1, 0.2]), period=1)
    assert "Input ``rp_bins`` must be a monotonically increasing 1D array" in err.value.args[0]
    with pytest.raises(ValueError) as err:
        npairs_xy_z(d1, d2, np.array([0.1, 0.2]), np.array([0.1, 0.3, 0.2]), period=1)
    assert "Input ``pi_bins`` must be a monotonically increasing 1D array with at least two entries" in \
        err.value.args[0]
    with pytest.raises(ValueError) as err:
        npairs_projected(d1, d2, np.array([0.1, 0.2]), 0.5, period=1)
    assert "Input ``pi_max`` must be a positive scalar less than period[2]/3" in err.value.args[0]


def test_set_approximate_cell_sizes():
    # PC/test_pair_counters/test_mesh_helpers.py:10-29
    with pytest.raises(ValueError) as err:
        P.set_approximate_cell_sizes(0.1, 0.1, 1)
    assert "Input ``approx_cell1_size`` must be a length-3 sequence" in err.value.args[0]
    with pytest.raises(ValueError) as err:
        P.set_approximate_cell_sizes([0.1, 0.1, 0.1], 0.1, 1)
    assert "Input ``approx_cell2_size`` must be a length-3 sequence" in err.value.args[0]


def test_wfunc_signatures():
    # PC/test_pair_counters/test_marked_npairs_3d.py:181-232
    widths = [_func_signature_int_from_wfunc(i) for i in range(1, 18)]
    assert widths == [1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 4, 4, 3, 3, 5, 5]
    with pytest.raises(HalotoolsError):
        _func_signature_int_from_wfunc(18)
    with pytest.raises(HalotoolsError):
        _func_signature_int_from_wfunc(0)
    with pytest.raises(ValueError):
        _func_signature_int_from_wfunc(1.0)


@pytest.mark.parametrize("wid", list(range(1, 18)))
def test_marked_wrong_weight_width_raises(wid):
    rng = np.random.RandomState(fixed_seed)
    s = rng.random_sample((200, 3))
    width = _func_signature_int_from_wfunc(wid) + 1
    w = np.random.RandomState(fixed_seed).random_sample(200 * width).reshape(200, width) - 0.5
    with pytest.raises(HalotoolsError):
        marked_npairs_3d(s, s, np.array([0.0, 0.1, 0.2, 0.3]), wid, period=1.0, weights1=w, weights2=w)


def test_marked_xy_z_wfunc_signatures():
    # PC/test_pair_counters/test_marked_npairs_xy_z.py:194-248: ids 1-15, widths as for 3-D
    from halotools_amd.mock_observables.pair_counters.marked_npairs_xy_z import \
        _func_signature_int_from_wfunc as sig_xyz
    assert [sig_xyz(i) for i in range(1, 16)] == [1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 4, 4, 3, 3]
    for bad in (0, 16, 17):
        with pytest.raises(HalotoolsError) as err:
            sig_xyz(bad)
        assert "The value ``weight_func_id`` = %i is not recognized" % bad in err.value.args[0]
    with pytest.raises(ValueError):
        sig_xyz(1.0)


@pytest.mark.parametrize("wid", list(range(1, 16)))
def test_marked_xy_z_wrong_weight_width_raises(wid):
    rng = np.random.RandomState(fixed_seed)
    s = rng.random_sample((200, 3))
    width = _func_signature_int_from_wfunc(wid) + 1
    w = np.random.RandomState(fixed_seed).random_sample(200 * width).reshape(200, width) - 0.5
    with pytest.raises(HalotoolsError):
        marked_npairs_xy_z(s, s, np.array([0.0, 0.1, 0.2, 0.3]), np.array([0, 0.15]), period=1.0, weights1=w,
                           weights2=w, weight_func_id=wid)


def test_marked_xy_z_default_wfunc_rejected():
    # the reference's default weight_func_id=0 is not a marking function of the xy_z engine
    s = np.random.RandomState(2).random_sample((50, 3))
    with pytest.raises(HalotoolsError):
        marked_npairs_xy_z(s, s, np.array([0.1, 0.2]), np.array([0.1, 0.2]), period=1.0)


def _request(kind, flags=0, wfunc=1, nw=1, nr=4, np_=2):
    mock = E.HpMock()
    mock.n1, mock.n2 = 100, 100
    mock.period[:] = [1.0, 1.0, 1.0]
    mock.pbc = 1
    r2 = (ctypes.c_double * nr)(*[(0.1 * (i + 1)) ** 2 / 4 for i in range(nr)])
    p2 = (ctypes.c_double * np_)(*[(0.1 * (i + 1)) ** 2 for i in range(np_)])
    req = E.HpRequest()
    req.kind, req.flags, req.nmocks, req.wfunc, req.nw, req.nr, req.np = kind, flags, 1, wfunc, nw, nr, np_
    req.mocks = ctypes.pointer(mock)
    req.r2 = ctypes.cast(r2, ctypes.POINTER(ctypes.c_double))
    req.p2 = ctypes.cast(p2, ctypes.POINTER(ctypes.c_double))
    return req, (mock, r2, p2)


def test_marked_xy_z_request_validation():
    lib = E.library()
    req, keep = _request(E.HP_MARKED_NPAIRS_XY_Z)
    assert lib.hp_workspace_size(ctypes.byref(req)) > 0
    for flags in (E.HP_ROWS1, E.HP_ALSO_COUNTS, E.HP_TANGENT, E.HP_NO_SUMS):
        req, keep = _request(E.HP_MARKED_NPAIRS_XY_Z, flags=flags)
        assert lib.hp_workspace_size(ctypes.byref(req)) == 0
        assert "takes no flags" in lib.hp_last_error().decode()
    req, keep = _request(E.HP_MARKED_NPAIRS_XY_Z, wfunc=16, nw=5)
    assert lib.hp_workspace_size(ctypes.byref(req)) == 0
    assert "1..15" in lib.hp_last_error().decode()
    # any number of bins: wide histograms are counted over several launches
    for nr, np_, ok in ((100, 100, True), (1 << 14, 1 << 13, False)):
        req, keep = _request(E.HP_MARKED_NPAIRS_XY_Z, nr=nr, np_=np_)
        r2 = np.linspace(1e-4, 0.01, nr)
        p2 = np.linspace(0.0, 0.01, np_)
        req.r2 = r2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        req.p2 = p2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        size = lib.hp_workspace_size(ctypes.byref(req))
        assert (size > 0) == ok, lib.hp_last_error()
    assert "2^26" in lib.hp_last_error().decode()


def test_marked_weight_length_message():
    s = np.random.RandomState(1).random_sample((50, 3))
    with pytest.raises(HalotoolsError) as err:
        marked_npairs_3d(s, s, np.array([0.1, 0.2]), 1, period=1.0, weights1=np.ones(49))
    assert "does not have the correct length" in err.value.args[0]
    with pytest.raises(HalotoolsError) as err:
        marked_npairs_3d(s, s, np.array([0.1, 0.2]), 1, period=1.0, weights1=np.ones((50, 1, 1)))
    assert "You must either pass in a 1-D or 2-D array" in err.value.args[0]


@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_hilbert_order_is_face_adjacent_bijection(bits):
    lib = E.library()
    n = 1 << bits
    idx = {}
    for x in range(n):
        for y in range(n):
            for z in range(n):
                idx[lib.hp_debug_hilbert(x, y, z, bits)] = (x, y, z)
    assert sorted(idx) == list(range(n ** 3))
    for h in range(n ** 3 - 1):
        a, b = np.array(idx[h]), np.array(idx[h + 1])
        assert np.abs(a - b).sum() == 1, (h, idx[h], idx[h + 1])


def test_mesh_dependent_bins_rejected():
    d1, d2 = _loci()
    with pytest.raises(ValueError):
        npairs_3d(d1, d2, np.array([0.1, -0.2]), period=1)


def test_segment_sum_scratch_and_validation():
    lib = E.library()
    offs = (ctypes.c_int64 * 4)(0, 4096, 4097, 20000)
    assert lib.hp_segment_sum_scratch(offs, 3) == 1 + 1 + 4  # ceil(n / 4096) chunks per mock
    bad = (ctypes.c_int64 * 3)(0, 10, 5)
    assert lib.hp_segment_sum_scratch(bad, 2) < 0
    assert "non-decreasing" in lib.hp_last_error().decode()
    out = ctypes.c_double()
    assert lib.hp_segment_sum(None, offs, 3, ctypes.byref(out), None, 0, None) < 0


def test_xyz_rows_detects_one_block_of_rows():
    # NumPy (N, 3) samples go to the device as one copy when their columns are
    # the 24-byte-strided views of one C-ordered float64 block
    from halotools_amd import _device as D
    s = np.random.RandomState(3).random_sample((1000, 3))
    rows = D.xyz_rows(P.columns(s))
    assert rows is not None and rows.shape == (1000, 3) and np.array_equal(rows, s)
    assert np.shares_memory(rows, s)
    # anything else takes the column-by-column path
    assert D.xyz_rows(P.columns(np.asfortranarray(s))) is None
    assert D.xyz_rows(P.columns(s.astype(np.float32))) is None
    assert D.xyz_rows(P.columns(np.random.random_sample((10, 4)))) is None
    assert D.xyz_rows((s[:, 1], s[:, 0], s[:, 2])) is None
    assert D.xyz_rows(P.columns(s[::2])) is None
    assert D.xyz_rows(P.columns(np.empty((0, 3)))) is None
    ro = s.copy()
    ro.flags.writeable = False
    assert D.xyz_rows(P.columns(ro)) is None


def test_tensor_xyz_rows_detects_one_block_of_rows():
    import torch
    from halotools_amd import _device as D
    t = torch.from_numpy(np.random.RandomState(4).random_sample((500, 3)))
    rows = D.tensor_xyz_rows(P.columns(t))
    assert rows is not None and tuple(rows.shape) == (500, 3) and torch.equal(rows, t)
    assert rows.data_ptr() == t.data_ptr()
    # the request reads such columns in place with a stride of 3 doubles
    cols, stride = D._coordinate_columns(P.columns(t))
    assert stride == 3 and [c.data_ptr() for c in cols] == [t.data_ptr() + 8 * k for k in range(3)]
    x, y, z = D._row_columns(rows)
    assert torch.equal(x, t[:, 0]) and torch.equal(y, t[:, 1]) and torch.equal(z, t[:, 2])
    # contiguous columns keep stride 1; other layouts are copied to contiguous columns
    cols, stride = D._coordinate_columns(tuple(c.contiguous() for c in (x, y, z)))
    assert stride == 1
    cols, stride = D._coordinate_columns((t[:, 2], t[:, 1], t[:, 0]))
    assert stride == 1 and all(c.is_contiguous() for c in cols) and torch.equal(cols[0], t[:, 2])
    assert D.tensor_xyz_rows(P.columns(t.float())) is None
    assert D.tensor_xyz_rows(P.columns(t[::2])) is None
    assert D.tensor_xyz_rows((t[:, 2], t[:, 1], t[:, 0])) is None
    assert D.tensor_xyz_rows(P.columns(torch.zeros((10, 4), dtype=torch.float64))) is None
    assert D.tensor_xyz_rows(P.columns(t.t().contiguous().t())) is None
    assert D.tensor_xyz_rows(P.columns(torch.zeros((0, 3), dtype=torch.float64))) is None


def test_request_planning_sweep():
    """hp_workspace_size over seeded random requests: every kind and flag
    combination, several mocks of uneven size (some empty), periodic or not,
    cell hints, edge counts up to the table limits, and invalid values.  Each
    request is planned twice (the size is a pure function of the request);
    an invalid one returns 0 with a message.  The entry points themselves
    reject the same requests without a workspace, before touching any device
    pointer.  (Run under AddressSanitizer / UBSan by test_sanitizers.py.)"""
    lib = E.library()
    rng = np.random.default_rng(2024)
    entries = [lib.hp_npairs_3d, lib.hp_npairs_projected, lib.hp_npairs_xy_z, lib.hp_marked_npairs_3d,
               lib.hp_npairs_per_object_3d, lib.hp_marked_npairs_xy_z, lib.hp_npairs_jackknife_3d,
               lib.hp_npairs_jackknife_xy_z]
    planned = rejected = 0
    for it in range(400):
        kind = int(rng.integers(0, 8))
        M = int(rng.integers(1, 5))
        mocks = (E.HpMock * M)()
        for m in range(M):
            n = int(rng.choice([0, 1, 63, 64, 65, 1000, int(rng.integers(0, 200000))]))
            mocks[m].n1, mocks[m].n2 = n, int(rng.integers(0, 5000))
            L = float(rng.uniform(50, 2000))
            mocks[m].period[:] = [L, L * rng.uniform(0.5, 2), L * rng.uniform(0.5, 2)]
            mocks[m].pbc = int(rng.integers(0, 2))
            if rng.random() < 0.3:
                mocks[m].cell[:] = [float(rng.uniform(-1, 60))] * 3
        nr = int(rng.choice([1, 2, 3, 20, 64, 65, 128, 129, 511, 512, 513, 3000]))
        np_ = int(rng.choice([0, 1, 2, 41, 100, 512, 513, 5000]))
        r2 = np.sort(rng.uniform(0.0, 15.0, nr)) ** 2
        if rng.random() < 0.1:
            r2 = r2[::-1].copy()        # non-monotone edges are accepted (suffix minimum)
        if rng.random() < 0.05:
            r2[0] = np.nan
        p2 = np.sort(rng.uniform(0.0, 15.0, max(np_, 1))) ** 2
        req = E.HpRequest()
        req.kind = kind
        req.flags = int(rng.integers(0, 64))
        req.nmocks = M
        req.wfunc = int(rng.integers(0, 19))
        req.nw = int(rng.integers(0, 7))
        req.nr = nr
        req.np = np_
        req.n_samples = int(rng.integers(0, 300))
        req.tile_waves = int(rng.choice([0, 1, 2, 3, 4, 8]))
        req.mocks = ctypes.cast(mocks, ctypes.POINTER(E.HpMock))
        req.r2 = r2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        req.p2 = p2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        size = lib.hp_workspace_size(ctypes.byref(req))
        assert size == lib.hp_workspace_size(ctypes.byref(req))
        if size == 0:
            rejected += 1
            assert lib.hp_last_error().decode()
        else:
            planned += 1
            assert size % 256 == 0
        # no workspace: rejected before any device work
        assert entries[kind](ctypes.byref(req)) < 0
        assert lib.hp_last_error().decode()
    assert planned > 50 and rejected > 50, (planned, rejected)


def test_noalloc_entry_rejects_before_device_work():
    # argument and size errors of hp_npairs_3d_noalloc come from host planning alone
    lib = E.library()
    req, keep = _request(E.HP_NPAIRS_3D)
    req.workspace = 16
    req.workspace_bytes = 16
    assert lib.hp_npairs_3d_noalloc(ctypes.byref(req)) < 0
    assert "takes no workspace" in lib.hp_last_error().decode()
    req, keep = _request(E.HP_NPAIRS_3D)
    keep[0].n1 = keep[0].n2 = 5000000
    assert lib.hp_workspace_size(ctypes.byref(req)) > (64 << 20)
    assert lib.hp_npairs_3d_noalloc(ctypes.byref(req)) < 0
    assert "more than the library's fixed 67108864 bytes" in lib.hp_last_error().decode()
    assert lib.hp_npairs_3d_noalloc(None) < 0


def test_noalloc_capacity_matches_documentation():
    """The no-alloc entry's fixed scratch holds an auto-count of about 720K
    points with 20 bins (the figure the header and the notes quote)."""
    lib = E.library()

    def size(n):
        mock = E.HpMock(n1=n, n2=n, pbc=1)
        mock.period[:] = [1000.0] * 3
        r2 = np.ascontiguousarray(np.logspace(-1, np.log10(50), 20) ** 2)
        req = E.HpRequest(kind=E.HP_NPAIRS_3D, nmocks=1, nr=20, flags=E.HP_SAME_SAMPLES)
        req.mocks = ctypes.pointer(mock)
        req.r2 = r2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return lib.hp_workspace_size(ctypes.byref(req))

    assert 0 < size(700000) <= (64 << 20) < size(760000)
