"""torch.autograd wrapper of the differentiable weighted PageRank.

Test harness only: the product is the C-ABI entry gbx_pr_adjoint; this
wrapper lets a test write ``loss.backward()`` over it.  Forward: the weighted
stochastic pull PageRank p(W, t) on the GPU.  Backward: gbx_pr_adjoint with
the incoming gradient as the loss weights gives dL/dW (forward-CSR order,
mapped back to the caller's edge order) and dL/dt = (1 - d) lambda.
"""
from __future__ import annotations

import numpy as np
import torch

import graphblox_amd as gbx


def edge_order(src, dst, fwd_src, fwd_dst, N):
    """Index into the forward CSR of every edge of an (src, dst) list."""
    key_fwd = fwd_src.astype(np.uint64) * np.uint64(N) + fwd_dst.astype(np.uint64)
    key = np.asarray(src, np.uint64) * np.uint64(N) + np.asarray(dst, np.uint64)
    order = np.argsort(key_fwd)
    pos = np.searchsorted(key_fwd[order], key)
    assert np.array_equal(key_fwd[order][pos], key), "edge missing from the device graph"
    return order[pos]


class WeightedPageRank(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, t, gctx, src, dst, N, iterations, epsilon):
        g = gctx.graph(src, dst, w.detach().cpu().numpy().astype(np.float32), num_vertices=N,
                       flags=gbx.GRAPH_WEIGHTED)
        f = g.export()
        tt = None if t is None else t.detach().cpu().numpy().astype(np.float32)
        ranks, rep = g.pagerank(iterations=iterations, epsilon=epsilon, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1,
                                weighted=True, teleport=tt)
        ctx.graph, ctx.tele = g, tt
        ctx.perm = edge_order(src, dst, f["src"], f["dst"], N)
        ctx.opts = (iterations, epsilon)
        ctx.has_t = t is not None
        return torch.from_numpy(ranks)

    @staticmethod
    def backward(ctx, grad_p):
        it, eps = ctx.opts
        out = ctx.graph.adjoint(grad_p.detach().cpu().numpy().astype(np.float32), teleport=ctx.tele,
                                iterations=it, epsilon=eps)
        ctx.graph.free()
        gw = torch.from_numpy(out["grad_w"][ctx.perm])
        gt = torch.from_numpy(out["grad_t"]) if ctx.has_t else None
        return gw, gt, None, None, None, None, None, None


def weighted_pagerank(w, t, gctx, src, dst, N, iterations=300, epsilon=1e-7):
    return WeightedPageRank.apply(w, t, gctx, src, dst, N, iterations, epsilon)
