"""Many mock realisations of uneven size counted in one launch per pass.

The reference counts one sample at a time: every call builds its own mesh and
runs its engine over it (PC/npairs_3d.py:111-150; PC =
halotools/mock_observables/pair_counters), and model fitting calls it once per
mock (halotools/empirical_models/factories/mock_factory_template.py:288-320).
Here all mocks are concatenated: one segmented counting sort builds every
mock's cells (keys local to a mock, sort tiles that never cross one), one
pair-kernel launch covers the lane groups of all mocks (each group knows its
mock's geometry), and the per-mock outputs come out of one fixed-order
reduction.

``HODBatchStep`` is the model-fitting step of the benchmark: for every mock,
the npairs_3d auto-counts, the product-weighted marked sums with Zheng07
central-occupation weights, and the gradient of a loss sum_k c_k W_k with
respect to each mock's logMmin, all capturable in one HIP graph: by default
one forward-mode pair pass carrying (w, dw/dlogMmin) per point; or, in
reverse mode, a count pass, a marked pass producing the Jacobian rows, the
backward contraction and a fixed-order per-mock sum.
"""
import math

import numpy as np
import torch

from . import _device as D
from . import _engine as E
from . import autotune
from .autodiff import _MarkedProductSums, suffix_sums


class MockBatch(object):
    """Concatenated realisations sharing one periodic box size and one set of bins.

    points : list of (N_m, 3) arrays/tensors, or one (sum N_m, 3) array with `sizes`
    period : box length (scalar or 3-vector), the same for every mock
    rbins  : radial bin edges (validated like npairs_3d's rbins)
    cells  : optional cell edge per mock (else the tuner's cache, else the library default)
    tile_waves : optional wavefronts per pair-kernel workgroup for every launch
             (else the tuned shape of each launch's kind, else the library's choice)
    """

    def __init__(self, points, period, rbins, sizes=None, device=None, cells=None, tile_waves=0):
        from .mock_observables.pair_counters import _process as P
        self.device = D.require_gpu(device)
        if sizes is None:
            sizes = [len(p) for p in points]
            if all(isinstance(p, torch.Tensor) for p in points):
                allp = torch.cat([p.to(self.device, torch.float64) for p in points])
            else:
                allp = torch.from_numpy(np.concatenate([np.asarray(p, dtype=np.float64) for p in points]))
        else:
            allp = points if isinstance(points, torch.Tensor) else torch.from_numpy(np.asarray(points, np.float64))
        allp = allp.to(self.device, torch.float64)
        self.sizes = [int(n) for n in sizes]
        self.M = len(self.sizes)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        # the column views of one (N, 3) block: the library reads the rows in place
        # (one 24-byte span per point in the key pass and the lane reads)
        self.rows = allp.contiguous()
        self.cols = D.RowColumns(self.rows.unbind(1))
        self.rbins = P.process_bins(rbins, P.RBINS_MSG)
        self.period = P.process_period(period)
        rmax = float(np.max(self.rbins))
        P.check_search_lengths((rmax, rmax, rmax), tuple(self.period))
        self.rmax = rmax
        # the caller's workgroup shape for every launch (0: per kind, see waves_for)
        self.tile_waves = int(tile_waves)
        self._kind_waves = {}  # tuned workgroup shape of the heaviest mock, per tuner kind
        self.mocks = [dict(n1=n, n2=n, period=list(self.period), pbc=True, cell=[0.0] * 3) for n in self.sizes]
        if cells is not None:
            for m in range(self.M):
                self.mocks[m]["cell"] = [float(cells[m])] * 3
        else:
            self._apply_tuning(E.HP_NPAIRS_3D)
        mock_id = np.repeat(np.arange(self.M), self.sizes)
        self.mock_id = torch.from_numpy(mock_id).to(self.device)

    def _apply_tuning(self, kind):
        # per-mock cell edges from the tuner's cache; the workgroup shape (one
        # per launch) from the entry of the mock with the most pair work, kept
        # for launches of the tuned kind only (the shape that suits one kernel
        # mode can starve another: 8-wave workgroups of the per-lane Jacobian
        # rows leave one workgroup per CU)
        heaviest = int(np.argmax(self.sizes)) if self.M else 0
        for m, n in enumerate(self.sizes):
            hit = autotune.lookup(kind, n, self.period, self.rmax)
            if hit is not None:
                self.mocks[m]["cell"] = [hit[0]] * 3
                if m == heaviest:
                    self._kind_waves[kind] = hit[1]

    def waves_for(self, kind):
        """Workgroup waves for launches of tuner kind `kind` (E.HP_NPAIRS_3D,
        E.HP_MARKED_NPAIRS_3D or autotune.TANGENT): the caller's tile_waves when
        given, else the tuned shape of that kind, else 0 (the library's choice)."""
        if self.tile_waves:
            return self.tile_waves
        return self._kind_waves.get(kind, 0)

    def tune(self, kind=autotune.TANGENT, **kw):
        """Autotune cell edges and the tile shape on this batch's own mocks.

        One representative (the largest mock) per density bucket of the tuner's
        cache key is timed; every mock of the bucket then uses its result.
        Returns {bucket key: ((target, waves), timings)}.
        """
        done = {}
        for m in np.argsort(self.sizes)[::-1]:
            n = self.sizes[m]
            k = autotune.key(kind, n, self.period, self.rmax)
            if k in done or n < 64:
                continue
            a, b = int(self.offsets[m]), int(self.offsets[m + 1])
            pts = self.rows[a:b]
            done[k] = autotune.tune(pts, self.period, self.rbins, kind=kind, device=self.device, **kw)
        self._apply_tuning(kind)
        return done

    def npairs_3d(self, build_events=None, pair_events=None):
        """(M, K) int64 auto-counts, one row per mock (npairs_3d(p_m, p_m, rbins, period)).

        build_events / pair_events: optional (start, end) HIP event handles
        recorded around the cell build and the pair kernel (profiling).
        """
        res = D.run(E.HP_NPAIRS_3D, self.mocks, self.cols, self.cols, self.rbins * self.rbins, same=True,
                    device=self.device, tile_waves=self.waves_for(E.HP_NPAIRS_3D), build_events=build_events,
                    pair_events=pair_events)
        self._check(res)
        return res.counts

    def marked_npairs_3d_product(self, weights):
        """(M, K) product-weighted auto sums, differentiable in `weights` (one weight per point)."""
        w = weights.to(self.device, torch.float64)
        plan = dict(mocks=self.mocks, s1=self.cols, s2=self.cols, r2=self.rbins * self.rbins, device=self.device,
                    same_weights=True, offsets1=list(self.offsets), offsets2=list(self.offsets),
                    tile_waves=self.waves_for(E.HP_MARKED_NPAIRS_3D))
        return _MarkedProductSums.apply(w, w, plan)

    @staticmethod
    def _check(res):
        if int(res.cbuf[-1].item()) & 0xFFFFFFFF:
            raise ValueError("All points must be finite and lie inside the periodic box [0, period]")


class HODBatchStep(object):
    """Counts + marked sums + their gradient for every mock of a batch, graph-capturable.

    Weights are Zheng07 central occupations w_i = 0.5 (1 + erf((log10 M_i - logMmin_m) / sigma)),
    restated from halotools/empirical_models/occupation_models/zheng07_components.py:149-153,
    with one logMmin per mock.  The loss of mock m is sum_k c_k W_mk; the step
    produces counts (M, K), W (M, K), loss (M,) and dloss/dlogMmin (M,).

    mode="tangent" (default): one symmetric pair pass carries (w, dw/dlogMmin)
    per point and accumulates the counts, W and dW/dlogMmin together
    (forward-mode: one parameter per mock, so no per-point rows are needed).
    mode="reverse": a count pass, a marked pass writing the per-point Jacobian
    rows, the hand-written backward contraction and a fixed-order per-mock sum
    -- the same arithmetic as the autograd path of ``marked_npairs_3d_product``.
    """

    def __init__(self, batch, log_mass, logMmin, sigma, c, mode="tangent"):
        if mode not in ("tangent", "reverse"):
            raise ValueError("mode must be 'tangent' or 'reverse'")
        self.mode = mode
        self.b = batch
        dev = batch.device
        self.log_mass = log_mass.to(dev, torch.float64).contiguous()
        self.logMmin = logMmin.to(dev, torch.float64).clone()
        self.sigma = float(sigma)
        self.c = c.to(dev, torch.float64)
        n = int(batch.offsets[-1])
        r2 = batch.rbins * batch.rbins
        self.grad_logMmin = torch.zeros(batch.M, dtype=torch.float64, device=dev)
        self.loss = torch.zeros(batch.M, dtype=torch.float64, device=dev)
        if mode == "tangent":
            self.wt = torch.empty((n, 2), dtype=torch.float64, device=dev)
            self.res = D.prepare(E.HP_MARKED_NPAIRS_3D, batch.mocks, batch.cols, batch.cols, r2, same=True,
                                 w1=self.wt, nw=2, wfunc=1, want_counts=False, want_sums=True, also_counts=True,
                                 tangent=True, device=dev, tile_waves=batch.waves_for(autotune.TANGENT))
            self.res_counts = self.res
        else:
            self.w = torch.empty((n, 1), dtype=torch.float64, device=dev)
            self.grad_w1 = torch.empty(n, dtype=torch.float64, device=dev)
            self.grad_w2 = torch.empty(n, dtype=torch.float64, device=dev)
            self.G = suffix_sums(self.c).expand(batch.M, -1).contiguous()
            # counts from the (symmetric) count pass; the marked pass keeps only
            # its FP64 histograms in LDS, which leaves room for more waves per CU
            self.res_counts = D.prepare(E.HP_NPAIRS_3D, batch.mocks, batch.cols, batch.cols, r2, same=True,
                                        device=dev, tile_waves=batch.waves_for(E.HP_NPAIRS_3D))
            self.res = D.prepare(E.HP_MARKED_NPAIRS_3D, batch.mocks, batch.cols, batch.cols, r2, same=True,
                                 w1=self.w, nw=1, wfunc=1, want_counts=False, want_sums=True, rows1=True,
                                 rows2=True, device=dev, tile_waves=batch.waves_for(E.HP_MARKED_NPAIRS_3D))
            self.dtheta = torch.empty(n, dtype=torch.float64, device=dev)
            self.segsum = D.SegmentSum(batch.offsets, dev)
        self.graph = None

    def _body(self):
        b = self.b
        if self.mode == "tangent":
            # weights and dw/dlogMmin in one library kernel (the same values as
            # the torch expression of the reverse branch below)
            D.zheng07_tangent(self.log_mass, b.mock_id, self.logMmin, self.sigma, self.wt)
            D.launch(self.res)
            torch.sum(self.res.tangents * self.c, dim=1, out=self.grad_logMmin)
        else:
            u = (self.log_mass - self.logMmin[b.mock_id]) / self.sigma
            w = 0.5 * (1.0 + torch.erf(u))
            dw = -torch.exp(-u * u) / (math.sqrt(math.pi) * self.sigma)   # dw / dlogMmin
            self.w[:, 0].copy_(w)
            D.launch(self.res_counts)
            D.launch(self.res)
            D.backward_rows(self.res.rows1, self.G, b.offsets, out=self.grad_w1)
            D.backward_rows(self.res.rows2, self.G, b.offsets, out=self.grad_w2)
            torch.mul(self.grad_w1 + self.grad_w2, dw, out=self.dtheta)
            self.segsum(self.dtheta, self.grad_logMmin)
        # outputs are written in place, so replays and eager steps share them
        torch.sum(self.res.sums * self.c, dim=1, out=self.loss)

    def run(self):
        """One eager step; returns (counts, W, loss, dloss/dlogMmin)."""
        self._body()
        return self.outputs()

    def capture(self):
        """Record the whole step into one HIP graph (replayed by ``replay``)."""
        s = torch.cuda.Stream(device=self.b.device)
        s.wait_stream(torch.cuda.current_stream(self.b.device))
        with torch.cuda.stream(s):
            self._body()  # warm-up outside capture
        torch.cuda.current_stream(self.b.device).wait_stream(s)
        torch.cuda.synchronize(self.b.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()
        return self

    def replay(self):
        self.graph.replay()
        return self.outputs()

    def outputs(self):
        return self.res_counts.counts, self.res.sums, self.loss, self.grad_logMmin

    def status(self):
        return (int(self.res.cbuf[-1].item()) | int(self.res_counts.cbuf[-1].item())) & 0xFFFFFFFF


def one_at_a_time(points, log_mass, logMmin, sigma, rbins, c, period):
    """The same work mock by mock through the drop-in API (the comparison path).

    Returns lists of counts, W, loss and dloss/dlogMmin per mock.
    """
    from .autodiff import marked_npairs_3d_product
    from .mock_observables import npairs_3d
    from .mocks import zheng07_ncen
    out = ([], [], [], [])
    for m in range(len(points)):
        p = points[m]
        counts = npairs_3d(p, p, rbins, period=period)
        th = logMmin[m].detach().clone().requires_grad_(True)
        w = zheng07_ncen(log_mass[m], th, sigma)
        W = marked_npairs_3d_product(p, p, rbins, w, w, period=period)
        loss = (W * c).sum()
        loss.backward()
        for lst, v in zip(out, (counts, W.detach(), loss.detach(), th.grad)):
            lst.append(v)
    return out
