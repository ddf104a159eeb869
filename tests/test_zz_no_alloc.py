"""hp_npairs_3d_noalloc (the npairs_3d entry point that allocates nothing)
against hp_npairs_3d with a caller workspace, through the C ABI: two calls in
a row on one GPU, the second with a smaller input (the fixed scratch keeps
the first call's contents), in auto and cross form; counts are bit-identical."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(res, gpu):
    import torch
    from halotools_amd import _engine as E
    req = E.HpRequest.from_buffer_copy(res.req)
    req.workspace = None
    req.workspace_bytes = 0
    req.stream = torch.cuda.current_stream(gpu).cuda_stream
    E.check(E.library().hp_npairs_3d_noalloc(ctypes.byref(req)))


def test_noalloc_matches_workspace_entry_twice(gpu):
    import torch
    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    rb = np.logspace(-1, np.log10(20), 12)
    rng = np.random.default_rng(77)
    for n1, n2, same in ((400000, 400000, True), (30000, 20000, False), (5000, 5000, True)):
        p1 = torch.tensor(rng.uniform(0, 300.0, (n1, 3)), device=gpu)
        p2 = p1 if same else torch.tensor(rng.uniform(0, 300.0, (n2, 3)), device=gpu)
        s1 = tuple(p1[:, i].contiguous() for i in range(3))
        s2 = s1 if same else tuple(p2[:, i].contiguous() for i in range(3))
        mock = [dict(n1=n1, n2=n2, period=[300.0] * 3, pbc=True)]
        ref = D.run(E.HP_NPAIRS_3D, mock, s1, s2, rb * rb, same=same, device=gpu)
        got = D.prepare(E.HP_NPAIRS_3D, mock, s1, s2, rb * rb, same=same, device=gpu)
        got.cbuf.fill_(-1)
        _noalloc(got, gpu)
        torch.cuda.synchronize()
        assert int(got.status.item()) & 0xFFFFFFFF == 0
        assert torch.equal(got.counts, ref.counts), (n1, n2, same)
        assert int(ref.counts[0, -1].item()) > 0


def test_noalloc_rejects_requests_beyond_its_scratch(gpu):
    from halotools_amd import _device as D
    from halotools_amd import _engine as E
    import torch
    n = 3000000
    x = torch.zeros(n, dtype=torch.float64, device=gpu)
    res = D.prepare(E.HP_NPAIRS_3D, [dict(n1=n, n2=n, period=[1000.0] * 3, pbc=True)], (x, x, x), (x, x, x),
                    np.array([1.0, 4.0]), same=True, device=gpu)
    with pytest.raises(E.EngineError, match="more than the library's fixed"):
        _noalloc(res, gpu)
