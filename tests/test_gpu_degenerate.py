"""Degenerate and ragged graphs through every PageRank / SpMV form: no edges
at all (one vertex; six vertices whose only edges were self loops removed by
dedupe), a star into one vertex (one row of every edge, every source
dangling but the centre's), a star out of one vertex (every row of one edge,
all but the centre dangling), and one row long enough to be split into hub
chunks.  Each result is checked against the oracle (the FP32 port of
pageRankPullGraphCSR / SPMVPullGraphCSR, the FP64 PageRank) and, for the
edgeless graphs, against the closed forms: reference semantics keep every
rank at base_pr = 1 - 0.85f (base_pr / N once normalised) and stop in the
first iteration; the stochastic form is uniform, 1/N."""
import ctypes

import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BASE_PR = np.float32(1.0) - np.float32(0.85)
TOL_F32 = 1e-5  # L1 against the reference port and FP64 (the FP32 parity bar; a long row's sum reassociates)
# BF16 contributions: each rounds with relative error <= 2^-9, a perturbation
# of at most 2^-9 of the propagated mass, which moves the stochastic fixed
# point by at most d/(1-d) times that in L1.  On these graphs the rounding
# errors do not cancel (the star's 63 equal contributions round alike), so
# the bound, not the R-MAT workloads' 2e-3, is the bar.
BF16_BOUND = 2.0 ** -9 * 0.85 / 0.15


def _star_in(n=64):
    s = np.arange(1, n, dtype=np.uint32)
    return s, np.zeros(n - 1, np.uint32)


def _star_out(n=64):
    d = np.arange(1, n, dtype=np.uint32)
    return np.zeros(n - 1, np.uint32), d


def _long_row(n=5001):
    # every vertex 1..n-1 into vertex 0, plus a chain 1 -> 2 -> ... -> n-1
    a = np.arange(1, n, dtype=np.uint32)
    s = np.concatenate([a, a[:-1]])
    d = np.concatenate([np.zeros(n - 1, np.uint32), a[1:]])
    return s, d


EDGED = {"star_in": _star_in, "star_out": _star_out, "long_row": _long_row}


def _oracle(s, d, N):
    fwd, inv = O.build_graph(s, d, None, N, mode=1)
    return fwd, inv, O.row_ptr_from(inv)


def _edgeless(ctx, which):
    if which == "empty":
        return ctx.graph(np.zeros(0, np.uint32), np.zeros(0, np.uint32)), np.zeros(0, np.uint32), 1
    s = np.array([3, 3, 5], np.uint32)
    return ctx.graph(s, s, flags=gbx.GRAPH_DEDUPE), np.zeros(0, np.uint32), 6


@pytest.mark.parametrize("which", ["empty", "self_loops_deduped"])
def test_edgeless_graphs(gpu_ctx, which):
    g, e, N = _edgeless(gpu_ctx, which)
    assert (g.N, g.E) == (N, 0)
    fwd, inv, rp = _oracle(e, e, N)
    # reference semantics: ranks stay at base_pr, so the first iteration has
    # no active vertex and the loop stops there (pageRank.c:338: 0 iterations)
    ranks, rep = g.pagerank(iterations=20, epsilon=1e-4)
    port = O.pr_pull_f32(fwd, inv, 20, 1e-4)
    assert l1(ranks, port["ranks"]) <= 1e-7
    np.testing.assert_allclose(ranks, float(BASE_PR) / N, rtol=1e-6)
    assert rep["iterations"] == port["iterations"] == 0
    # stochastic: every vertex dangling, the uniform distribution at every iteration
    rs, reps = g.pagerank(iterations=30, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE)
    r64, _, _ = O.pr_f64(N, fwd.out_degree, rp, inv.dst, mode=1, iterations=30)
    np.testing.assert_allclose(r64, 1.0 / N, rtol=1e-15)
    assert l1(rs, r64) <= 1e-6 and abs(reps["rank_sum"] - 1.0) < 1e-6
    rb, _ = g.pagerank(iterations=30, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE, precision=gbx.BF16)
    assert l1(rb, r64) <= BF16_BOUND
    # SpMV: no edge, y stays 0; x = 0.001 for dangling vertices (SPMV.c:156-163)
    x, y, _ = g.spmv(iterations=1)
    xo, yo = O.spmv_pull_f32(fwd, inv, 1)
    np.testing.assert_array_equal(y, yo)
    np.testing.assert_array_equal(x, xo)
    assert not y.any()
    # batched form: teleport to vertex 0 (the dangling mass still spreads
    # uniformly: (1-d) + d/N at 0, d/N elsewhere) and the uniform teleport
    s = g.ppr_session(2, [np.array([0]), np.arange(N)], iterations=30, epsilon=0.0, stop=gbx.STOP_NONE)
    s.reset()
    s.launch(30)
    rk, _ = s.read()
    s.free()
    t0 = np.zeros(N)
    t0[0] = 1.0
    r64t, _, _ = O.pr_f64(N, fwd.out_degree, rp, inv.dst, mode=1, iterations=30, teleport=t0)
    want = np.full(N, 0.85 / N)
    want[0] += 0.15
    np.testing.assert_allclose(r64t, want, rtol=1e-12)
    assert l1(rk[:, 0], r64t) <= 1e-6 and l1(rk[:, 1], r64) <= 1e-6
    # the binned method has nothing to bin: refused, the pull iteration runs these graphs
    with pytest.raises(gbx.GbxError):
        g.pagerank(iterations=5, epsilon=0.0, method=gbx.METHOD_BINNED)
    g.free()


def _hub_tuning(tuning):
    thr = (ctypes.c_uint32 * 7)(256, 256, 128, 64, 32, 16, 8)
    tuning(bin_thresholds=thr, hub_chunk=256)


@pytest.mark.parametrize("name", sorted(EDGED))
@pytest.mark.parametrize("method", [gbx.METHOD_PULL, gbx.METHOD_BINNED])
def test_edged_degenerate_graphs(gpu_ctx, tuning, name, method):
    if name == "long_row" and method == gbx.METHOD_PULL:
        _hub_tuning(tuning)  # the 5,000-edge row split into 20 hub chunks
    s, d = EDGED[name]()
    N = int(max(s.max(), d.max())) + 1
    g = gpu_ctx.graph(s, d, sort_mode=1)
    assert (g.N, g.E) == (N, s.size)
    fwd, inv, rp = _oracle(s, d, N)
    print(f"[gpu] degenerate {name} method {method}: N={N} E={s.size}", flush=True)
    ranks, rep = g.pagerank(iterations=20, epsilon=0.0, method=method)
    port = O.pr_pull_f32(fwd, inv, 20, 0.0)
    r64, _, _ = O.pr_f64(N, fwd.out_degree, rp, inv.dst, mode=0, iterations=20)
    print(f"[gpu]   REF L1 vs port {l1(ranks, port['ranks']):.3e}, vs FP64 {l1(ranks, r64):.3e} "
          f"(port vs FP64 {l1(port['ranks'], r64):.3e})", flush=True)
    assert l1(ranks, port["ranks"]) <= TOL_F32 and l1(ranks, r64) <= TOL_F32
    assert rep["executed"] == 20
    rs, reps = g.pagerank(iterations=40, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE, method=method)
    r64s, _, _ = O.pr_f64(N, fwd.out_degree, rp, inv.dst, mode=1, iterations=40)
    assert l1(rs, r64s) <= TOL_F32 and abs(reps["rank_sum"] - 1.0) < 1e-5
    rb, _ = g.pagerank(iterations=40, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE,
                       precision=gbx.BF16, method=method)
    print(f"[gpu]   BF16 stochastic L1 vs FP64 {l1(rb, r64s):.3e} (bound {BF16_BOUND:.3e})", flush=True)
    assert l1(rb, r64s) <= BF16_BOUND


@pytest.mark.parametrize("name", sorted(EDGED))
def test_edged_degenerate_spmv_and_batched(gpu_ctx, tuning, name):
    if name == "long_row":
        _hub_tuning(tuning)
    s, d = EDGED[name]()
    N = int(max(s.max(), d.max())) + 1
    g = gpu_ctx.graph(s, d, sort_mode=1)
    fwd, inv, rp = _oracle(s, d, N)
    x, y, _ = g.spmv(iterations=2)
    xo, yo = O.spmv_pull_f32(fwd, inv, 2)
    np.testing.assert_array_equal(x, xo)
    # y = 2 * sum over the row of 0.0001f * x[u], in FP64 from the same FP32
    # factors; the port's sequential FP32 sum of a row of n terms is within
    # 2n * 2^-24 of it, the GPU's (chunked) sum within 1e-6
    w = np.float64(np.float32(0.0001))
    rows = np.repeat(np.arange(N), np.diff(rp).astype(np.int64))
    y64 = 2.0 * np.bincount(rows, weights=w * xo[inv.dst].astype(np.float64),
                            minlength=N)
    nmax = int(np.diff(rp).max())
    np.testing.assert_allclose(yo, y64, rtol=2 * nmax * 2.0 ** -24, atol=0)
    np.testing.assert_allclose(y, y64, rtol=1e-6, atol=0)
    rng = np.random.default_rng(N)
    sets = [rng.choice(N, 3, replace=False), np.array([0]), np.arange(N)]
    ss = g.ppr_session(3, sets, iterations=40, epsilon=0.0, stop=gbx.STOP_NONE)
    ss.reset()
    ss.launch(40)
    rk, _ = ss.read()
    ss.free()
    for q, st in enumerate(sets):
        t = np.zeros(N)
        t[st] = 1.0 / len(st)
        r64, _, _ = O.pr_f64(N, fwd.out_degree, rp, inv.dst, mode=1, iterations=40, teleport=t)
        assert l1(rk[:, q], r64) <= TOL_F32, (q, l1(rk[:, q], r64))
