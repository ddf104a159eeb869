"""Differentiable product-weighted marked pair sums (weight_func_id = 1).

``marked_npairs_3d_product`` returns the cumulative sums of the reference's
``marked_npairs_3d(..., weight_func_id=1)`` (PC/marked_npairs_3d.py:25-183,
marking function w1*w2 of PC/marked_cpairs/marking_functions.pyx:14-21;
PC = halotools/mock_observables/pair_counters) as a torch tensor that carries
gradients to ``weights1`` and ``weights2`` -- and through them to whatever
produced the weights, e.g. HOD parameters of a mean-occupation model.
Positions, and therefore which bin each pair falls in, are exact and not
differentiated.

Forward.  One pass of the pair kernel in Jacobian mode writes, for every point
i of sample 1, the per-bin sums rows1[i, s] = sum of w2_j over the partners j
whose first bin is s (the derivative of the non-cumulative bin sums with
respect to w1_i).  A second pass with the roles swapped writes rows2 for the
sample-2 weights, with the pair arithmetic still (x1 - shift) - x2, so bin
membership is the forward's, bit for bit.  The sums are
W_k = sum_{s <= k} sum_i w1_i rows1[i, s].

Backward.  With upstream g = dL/dW and G(s) = sum_{k >= s} g_k, the hand-written
contraction kernel computes dL/dw1_i = sum_s rows1[i, s] G(s) (and the same for
rows2): one HBM-bound pass over the rows, no second pair traversal.
"""
import numpy as np
import torch

from . import _device as D
from . import _engine as E
from .mock_observables.pair_counters import _common as C
from .mock_observables.pair_counters import _process as P


def suffix_sums(g):
    """G[..., s] = sum_{k >= s} g[..., k] for upstream gradients g of cumulative sums."""
    return torch.flip(torch.cumsum(torch.flip(g, dims=[-1]), dim=-1), dims=[-1])


class _MarkedProductSums(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w1, w2, plan):
        need1 = ctx.needs_input_grad[0]
        need2 = ctx.needs_input_grad[1]
        same = plan["same_weights"]
        res = D.run(E.HP_MARKED_NPAIRS_3D, plan["mocks"], plan["s1"], plan["s2"], plan["r2"], same=same,
                    w1=w1.detach().reshape(-1, 1).contiguous(),
                    w2=None if same else w2.detach().reshape(-1, 1).contiguous(), nw=1, wfunc=1,
                    want_counts=False, want_sums=True, rows1=True, rows2=(need2 and not same) or same,
                    device=plan["device"], tile_waves=plan.get("tile_waves", 0))
        if int(res.cbuf[-1].item()) & 0xFFFFFFFF:
            raise ValueError("All points must be finite and lie inside the periodic box [0, period]")
        ctx.rows1 = res.rows1
        ctx.rows2 = res.rows2
        ctx.offsets1 = plan["offsets1"]
        ctx.offsets2 = plan["offsets2"]
        ctx.plan_res = res
        ctx.need = (need1, need2)
        return res.sums

    @staticmethod
    def backward(ctx, gsum):
        G = suffix_sums(gsum.to(torch.float64)).contiguous()
        g1 = g2 = None
        if ctx.need[0]:
            g1 = D.backward_rows(ctx.rows1, G, ctx.offsets1)
        if ctx.need[1]:
            g2 = D.backward_rows(ctx.rows2, G, ctx.offsets2)
        return g1, g2, None


def marked_npairs_3d_product(sample1, sample2, rbins, weights1, weights2, period=None,
                             num_threads=1, approx_cell1_size=None, approx_cell2_size=None):
    """Cumulative sums of w1_i * w2_j over pairs within each rbins edge, differentiable in the weights.

    Same arguments and validation as ``marked_npairs_3d(..., weight_func_id=1)``
    except that ``weights1`` / ``weights2`` are 1-D float64 torch tensors (any
    device; they may require grad).  Returns a float64 CUDA tensor of length
    ``len(rbins)``.  Passing the same tensor as both weights (an auto-count with
    one set of marks) gives the gradient through both factors of every pair.
    """
    num_threads = P.process_num_threads(num_threads)
    x1, y1, z1 = P.columns(sample1)
    x2, y2, z2 = P.columns(sample2)
    rbins = np.atleast_1d(rbins).astype("f8")
    rmax = np.max(rbins)
    rbins = P.process_bins(rbins, P.RBINS_MSG)
    P.check_last_edge(rbins)
    if period is None:
        pbc = False
        x1, y1, z1, x2, y2, z2, period = P.enclose_in_box(
            x1, y1, z1, x2, y2, z2, min_size=[rmax * 3.0, rmax * 3.0, rmax * 3.0])
    else:
        pbc = True
        period = P.process_period(period)
    a1, a2 = P.default_cell_sizes(approx_cell1_size, approx_cell2_size, (rmax, rmax, rmax))
    P.set_approximate_cell_sizes(a1, a2, period)
    P.check_search_lengths((rmax, rmax, rmax), tuple(period))
    n1, n2 = len(x1), len(x2)
    if tuple(weights1.shape) != (n1,) or tuple(weights2.shape) != (n2,):
        raise ValueError("weights1 and weights2 must be 1-D tensors with one weight per point")

    device = D.require_gpu(weights1.device if weights1.is_cuda else None)
    same_samples = D.same_array(sample1, sample2)
    s1, s2 = C.upload((x1, y1, z1), (x2, y2, z2), same_samples, device)
    w1 = weights1.to(device=device, dtype=torch.float64)
    same_weights = same_samples and (weights2 is weights1)
    w2 = w1 if same_weights else weights2.to(device=device, dtype=torch.float64)
    plan = dict(mocks=[C.mock_dict(n1, n2, period, pbc)], s1=s1, s2=s2, r2=P.squared(rbins), device=device,
                same_weights=same_weights, offsets1=[0, n1], offsets2=[0, n2])
    return _MarkedProductSums.apply(w1, w2, plan)[0]
