"""Differentiable weighted PageRank: the GPU adjoint (gbx_pr_adjoint, through
the torch.autograd wrapper of the test harness) against the FP64 CPU adjoint
of the oracle and against FP64 central finite differences, which use the
oracle's forward solve only.

Bars: dL/dW at every finite-difference edge within 1e-4 relative; the
directional derivative of the whole gradient within 1e-4 relative; p, lambda,
dL/dW and dL/dt within 1e-4 normwise of the FP64 adjoint."""
import math
import os

import numpy as np
import pytest
import torch

import graphblox_amd as gbx
from conftest import GOLDEN
from oracle import oracle as O
from pr_autograd import edge_order, weighted_pagerank

pytestmark = pytest.mark.gpu

SEED = 20261203
TOL = 1e-4


def _rel(a, b, ord=2):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b, ord) / np.linalg.norm(b, ord))


def _inv_csr(g):
    inv = g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    return inv, rp


def _small_graph(rng, n, m):
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    s, d, _ = O.dedupe_keep_first(src, dst)
    w = rng.lognormal(0.0, 1.0, s.size).astype(np.float32)
    return s, d, w


def test_adjoint_small_matches_fp64_and_finite_differences(gpu_ctx):
    rng = np.random.default_rng(SEED)
    n = 2000
    s, d, w = _small_graph(rng, n, 24000)
    N = int(max(s.max(), d.max())) + 1
    gl = rng.standard_normal(N)
    t = rng.random(N)
    t /= t.sum()
    print(f"[gpu] adjoint small N={N} E={s.size}", flush=True)
    g = gpu_ctx.graph(s, d, w, num_vertices=N, flags=gbx.GRAPH_WEIGHTED)
    out = g.adjoint(gl.astype(np.float32), teleport=t.astype(np.float32))
    assert out["fwd_report"]["converged"] and out["adj_report"]["converged"]
    inv, rp = _inv_csr(g)
    p, lam, gw, gt = O.adjoint_f64(N, rp, inv["dst"], inv["w"].astype(np.float64), gl.astype(np.float32),
                                   teleport=t.astype(np.float32), iterations=2000, tol=1e-15)
    f = g.export()
    perm = edge_order(inv["dst"], inv["src"], f["src"], f["dst"], N)  # in-CSR edge -> forward-CSR index
    gw_gpu = out["grad_w"][perm]
    errs = dict(p=_rel(out["p"], p, 1), lam=_rel(out["lam"], lam, 1), gw=_rel(gw_gpu, gw), gt=_rel(out["grad_t"], gt))
    print(f"[gpu] adjoint small vs FP64: {errs}", flush=True)
    assert max(errs.values()) <= TOL
    # central differences of the FP64 forward solve on 16 edges
    eligible = np.flatnonzero(g.export()["out_degree"][inv["dst"]] >= 2)
    edges = rng.choice(eligible, 16, replace=False)
    fd = O.fd_edge_grads(N, rp, inv["dst"], inv["w"].astype(np.float64), gl.astype(np.float32), edges, p,
                         h=2e-3, teleport=t.astype(np.float32), iterations=150)
    rel = np.abs(gw_gpu[edges] - fd) / np.abs(fd)
    print(f"[gpu] adjoint small vs FD: max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= TOL
    # dL/dt along zero-sum directions (the teleport stays a distribution)
    for _ in range(4):
        a, b = rng.choice(N, 2, replace=False)
        h = 1e-3 * min(t[a], t[b])
        tp, tm = t.copy(), t.copy()
        tp[a] += h
        tp[b] -= h
        tm[a] -= h
        tm[b] += h
        _, pp = O.loss_f64(N, rp, inv["dst"], inv["w"].astype(np.float64), gl.astype(np.float32), teleport=tp,
                           p0=p, iterations=400)
        _, pm = O.loss_f64(N, rp, inv["dst"], inv["w"].astype(np.float64), gl.astype(np.float32), teleport=tm,
                           p0=p, iterations=400)
        fd_t = math.fsum(gl.astype(np.float32).astype(np.float64) * (pp - pm)) / (2 * h)
        got = float(out["grad_t"][a]) - float(out["grad_t"][b])
        assert got == pytest.approx(fd_t, rel=TOL, abs=1e-7)


def test_adjoint_forward_equals_plain_weighted_pagerank(gpu_ctx):
    rng = np.random.default_rng(SEED + 1)
    s, d, w = _small_graph(rng, 5000, 80000)
    N = int(max(s.max(), d.max())) + 1
    g = gpu_ctx.graph(s, d, w, num_vertices=N, flags=gbx.GRAPH_WEIGHTED)
    out = g.adjoint(rng.standard_normal(N).astype(np.float32), iterations=200, epsilon=1e-7)
    ranks, rep = g.pagerank(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, weighted=True)
    assert rep["executed"] == out["fwd_report"]["executed"]
    assert np.array_equal(ranks, out["p"])


def test_adjoint_linear_in_loss_weights(gpu_ctx):
    # lambda and dL/dW are linear in g: g1 + 2 g2 against the two runs
    rng = np.random.default_rng(SEED + 2)
    s, d, w = _small_graph(rng, 3000, 40000)
    N = int(max(s.max(), d.max())) + 1
    g = gpu_ctx.graph(s, d, w, num_vertices=N, flags=gbx.GRAPH_WEIGHTED)
    g1 = rng.standard_normal(N).astype(np.float32)
    g2 = rng.standard_normal(N).astype(np.float32)
    kw = dict(iterations=400, epsilon=1e-7)
    a, b, c = g.adjoint(g1, **kw), g.adjoint(g2, **kw), g.adjoint(g1 + 2 * g2, **kw)
    assert _rel(c["lam"], a["lam"].astype(np.float64) + 2 * b["lam"]) <= 1e-5
    assert _rel(c["grad_w"], a["grad_w"].astype(np.float64) + 2 * b["grad_w"]) <= 1e-4


def test_adjoint_errors(gpu_ctx):
    rng = np.random.default_rng(SEED + 3)
    s, d, w = _small_graph(rng, 100, 400)
    N = int(max(s.max(), d.max())) + 1
    unweighted = gpu_ctx.graph(s, d, num_vertices=N)
    with pytest.raises(gbx.GbxError, match="weights"):
        unweighted.adjoint(np.ones(N, np.float32))
    weighted = gpu_ctx.graph(s, d, w, num_vertices=N, flags=gbx.GRAPH_WEIGHTED)
    batch = gpu_ctx.graph_batch([weighted, weighted])
    with pytest.raises(gbx.GbxError, match="batch"):
        batch.adjoint(np.ones(2 * N, np.float32))
    with pytest.raises(gbx.GbxError):
        weighted.adjoint(np.ones(N, np.float32), stop=gbx.STOP_ACTIVE)


def test_adjoint_rmat_s20_autograd(gpu_ctx):
    """The differentiable workload: R-MAT scale 20, edge factor 16, lognormal
    weights, loss weights N(0, 1); gradients through torch.autograd."""
    fx_path = os.path.join(GOLDEN, "adjoint_fd_s20.npz")
    with np.load(fx_path, allow_pickle=False) as z:
        fx = {k: z[k] for k in z.files}
    s, d, w, N, gl = O.weighted_rmat_workload(20, 16, SEED)
    assert (N, s.size) == (int(fx["N"]), int(fx["E"]))
    print(f"[gpu] adjoint R-MAT s20 N={N} E={s.size} (autograd)", flush=True)
    wt = torch.tensor(w, requires_grad=True)
    p = weighted_pagerank(wt, None, gpu_ctx, s, d, N, iterations=300, epsilon=1e-7)
    loss = (torch.from_numpy(gl.astype(np.float32)) * p).sum()
    loss.backward()
    gw = wt.grad.numpy().astype(np.float64)  # edge-list order
    assert float(loss.detach()) == pytest.approx(float(fx["loss"]), rel=1e-5)
    # 64 central finite differences (FP64 forward solves only)
    key = s.astype(np.uint64) * np.uint64(N) + d.astype(np.uint64)
    order = np.argsort(key)
    fk = fx["u"].astype(np.uint64) * np.uint64(N) + fx["v"].astype(np.uint64)
    at = order[np.searchsorted(key[order], fk)]
    assert np.array_equal(key[at], fk)
    rel = np.abs(gw[at] - fx["fd"]) / np.abs(fx["fd"])
    print(f"[gpu] s20 vs FD on 64 edges: max rel {rel.max():.2e}, median {np.median(rel):.2e}", flush=True)
    assert rel.max() <= TOL
    # directional derivative of the whole gradient (dot-product test)
    xi = np.random.default_rng(SEED + 400 + 20).standard_normal(s.size)
    dd = math.fsum(gw * xi * w.astype(np.float64))
    print(f"[gpu] s20 directional derivative {dd:.6e} vs FD {float(fx['dir_fd']):.6e}", flush=True)
    assert dd == pytest.approx(float(fx["dir_fd"]), rel=TOL)
    # the full FP64 CPU adjoint
    g = gpu_ctx.graph(s, d, w, num_vertices=N, flags=gbx.GRAPH_WEIGHTED)
    out = g.adjoint(gl.astype(np.float32), iterations=300, epsilon=1e-7)
    inv, rp = _inv_csr(g)
    p64, lam64, gw64, gt64 = O.adjoint_f64(N, rp, inv["dst"], inv["w"].astype(np.float64), gl.astype(np.float32),
                                           iterations=2000, tol=1e-15)
    perm = edge_order(inv["dst"], inv["src"], s, d, N)  # in-CSR edge -> edge-list index
    errs = dict(p=_rel(out["p"], p64, 1), lam=_rel(out["lam"], lam64, 1), gw=_rel(gw[perm], gw64),
                gt=_rel(out["grad_t"], gt64))
    print(f"[gpu] s20 vs FP64 adjoint: {errs}; iterations fwd {out['fwd_report']['executed']} "
          f"adj {out['adj_report']['executed']}", flush=True)
    assert max(errs.values()) <= TOL
