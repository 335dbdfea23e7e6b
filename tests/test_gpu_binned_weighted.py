"""The binned (propagation-blocking) iteration for the weighted forms: every
cold edge's weight rides beside its source index in the scatter and every
hot pair's beside the pair in the accumulate, so weighted stochastic
PageRank, the weighted SpMV (SPMV.c:167-194) and the adjoint fixed point
run without the pull's random gathers.

* weighted stochastic PageRank: binned against the FP64 oracle (row-
  normalised transition weights) and against the pull kernel, lognormal and
  uniform weights, personalised teleports, bit-reproducible run to run;
* weighted SpMV on the binned iteration (taken from binned_min_vertices on)
  against the port of SPMVPullGraphCSR over 1 and 3 accumulating passes,
  both directions;
* the adjoint on the binned iteration against the FP64 adjoint and the
  pull-method adjoint."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import l1
from oracle import oracle as O
from pr_autograd import edge_order

pytestmark = pytest.mark.gpu


def _oracle(g):
    f, i = g.export(False), g.export(True)
    mk = lambda c: O.OracleCSR(g.N, g.E, c["out_degree"], c["in_degree"], c["edges_idx"], c["src"], c["dst"], c["w"])
    fwd, inv = mk(f), mk(i)
    rp = O.row_ptr_from(inv)
    S = np.bincount(fwd.src, weights=fwd.w.astype(np.float64), minlength=g.N)
    wn = np.where(S[inv.dst] > 0, inv.w.astype(np.float64) / np.maximum(S[inv.dst], 1e-300), 0.0)
    return fwd, inv, rp, wn


def _weighted(ctx, scale, kind, seed):
    return ctx.rmat_graph(scale, 16, seed, weight_kind=kind, weight_seed=seed + 1, flags=gbx.GRAPH_WEIGHTED)


@pytest.mark.parametrize("scale,kind", [(16, 2), (18, 1), (20, 2)])
def test_weighted_pagerank_binned_vs_fp64_and_pull(gpu_ctx, scale, kind):
    g = _weighted(gpu_ctx, scale, kind, 100 + scale)
    fwd, inv, rp, wn = _oracle(g)
    kw = dict(iterations=200, epsilon=1e-8, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, weighted=True)
    rb, repb = g.pagerank(method=gbx.METHOD_BINNED, **kw)
    rb2, _ = g.pagerank(method=gbx.METHOD_BINNED, **kw)
    rp_, repp = g.pagerank(method=gbx.METHOD_PULL, **kw)
    r64, it64, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=200, epsilon=1e-8, weights=wn)
    eb, ep, ebp = l1(rb, r64), l1(rp_, r64), l1(rb, rp_)
    print(f"[gpu] weighted PR s{scale} weights {kind}: binned {repb['executed']} iterations (FP64 {it64}), "
          f"L1 vs FP64 binned {eb:.2e} pull {ep:.2e}, binned vs pull {ebp:.2e}", flush=True)
    np.testing.assert_array_equal(rb, rb2)  # exact row sums of a fixed layout
    assert repb["converged"] == 1 and abs(repb["executed"] - it64) <= 1
    assert eb <= 1e-5 and ebp <= 1e-5
    assert abs(repb["rank_sum"] - 1.0) < 1e-5
    g.free()


def test_weighted_pagerank_binned_personalised(gpu_ctx):
    g = _weighted(gpu_ctx, 17, 2, 7)
    fwd, inv, rp, wn = _oracle(g)
    rng = np.random.default_rng(7)
    t = np.zeros(g.N)
    t[rng.choice(g.N, 500, replace=False)] = 1.0 / 500
    r, _ = g.pagerank(iterations=100, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE, weighted=True,
                      teleport=t.astype(np.float32), method=gbx.METHOD_BINNED)
    r64, _, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, weights=wn, teleport=t)
    assert l1(r, r64) <= 1e-5
    g.free()


@pytest.mark.parametrize("direction", [gbx.PULL_IN, gbx.PULL_FWD])
def test_weighted_spmv_binned_vs_port(gpu_ctx, direction):
    saved = gpu_ctx.tuning()
    t = gpu_ctx.tuning()
    t.binned_min_vertices = 0  # the binned iteration for every size
    gpu_ctx.set_tuning(t)
    try:
        for scale, kind in ((16, 1), (19, 2)):
            g = _weighted(gpu_ctx, scale, kind, 40 + scale)
            fwd, inv, _, _ = _oracle(g)
            rows = inv if direction == gbx.PULL_IN else fwd
            for iters in (1, 3):
                x, y, sec = g.spmv(iterations=iters, weighted=True, direction=direction)
                xo, yo = O.spmv_pull_f32(fwd, rows, iterations=iters, weights=rows.w, nthreads=0)
                np.testing.assert_array_equal(x, xo)
                rel = l1(y, yo) / float(np.abs(yo.astype(np.float64)).sum())
                print(f"[gpu] binned weighted SpMV s{scale} direction {direction} passes {iters}: relative L1 vs "
                      f"port {rel:.2e} ({sec * 1e3:.3f} ms)", flush=True)
                np.testing.assert_allclose(y, yo, rtol=2e-5, atol=1e-9)
                assert rel <= 1e-6
            # a given input vector
            xi = np.random.default_rng(scale).random(g.N).astype(np.float32)
            _, y, _ = g.spmv(iterations=1, weighted=True, direction=direction, x=xi)
            yo = np.zeros(g.N, np.float64)
            np.add.at(yo, rows.src, rows.w.astype(np.float64) * xi[rows.dst].astype(np.float64))
            np.testing.assert_allclose(y, yo, rtol=2e-5, atol=1e-9)
            g.free()
    finally:
        gpu_ctx.set_tuning(saved)


def test_adjoint_binned_vs_fp64_and_pull(gpu_ctx):
    g = _weighted(gpu_ctx, 18, 2, 2018)
    rng = np.random.default_rng(18)
    gl = rng.standard_normal(g.N).astype(np.float32)
    ob = g.adjoint(gl, iterations=300, epsilon=1e-8, method=gbx.METHOD_BINNED)
    op = g.adjoint(gl, iterations=300, epsilon=1e-8, method=gbx.METHOD_PULL)
    inv = g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    p, lam, gw, gt = O.adjoint_f64(g.N, rp, inv["dst"], inv["w"].astype(np.float64), gl, iterations=2000, tol=1e-15)
    f = g.export()
    perm = edge_order(inv["dst"], inv["src"], f["src"], f["dst"], g.N)
    rel = lambda a, b, o=2: float(np.linalg.norm(np.asarray(a, np.float64) - b, o) / np.linalg.norm(b, o))
    errs = dict(p=rel(ob["p"], p, 1), lam=rel(ob["lam"], lam, 1), gw=rel(ob["grad_w"][perm], gw),
                gt=rel(ob["grad_t"], gt))
    vs_pull = dict(lam=rel(ob["lam"], op["lam"].astype(np.float64), 1),
                   gw=rel(ob["grad_w"], op["grad_w"].astype(np.float64)))
    print(f"[gpu] binned adjoint s18: {ob['fwd_report']['executed']} + {ob['adj_report']['executed']} iterations, "
          f"vs FP64 {errs}, vs pull {vs_pull}", flush=True)
    assert ob["fwd_report"]["converged"] and ob["adj_report"]["converged"]
    assert max(errs.values()) <= 1e-4 and max(vs_pull.values()) <= 1e-4
    g.free()


def test_weighted_binned_refuses_bf16(gpu_ctx):
    g = _weighted(gpu_ctx, 12, 1, 3)
    with pytest.raises(gbx.GbxError, match="binned"):
        g.pagerank(iterations=5, mode=gbx.PR_STOCHASTIC, weighted=True, precision=gbx.BF16,
                   method=gbx.METHOD_BINNED)
    g.free()
