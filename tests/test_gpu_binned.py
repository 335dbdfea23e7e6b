"""Binned (propagation-blocking) PageRank iteration: same parity bars as the
pull iteration (<= 1e-5 L1 vs the reference's FP32 ranks and vs FP64; BF16
<= 2e-3), deterministic (exact fixed-point row sums of a fixed layout), and
the same results under every layout cut (tiny chunks and bins force
thousands of segments and giant rows split over several accumulate blocks):
bit-identical with one V value per edge, within FP32 rounding of the run
sums (<= 1e-6 L1) with the run layout."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL, l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL]
TOL_F32 = 1e-5
BINNED = gbx.METHOD_BINNED


def _graph(ctx, z, **kw):
    return ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, **kw)


def _oracle_csr(z):
    return O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)


@pytest.fixture(params=[16, 64])
def small_cut_ctx(request):
    """A context whose binned layout uses tiny chunks and bins and only 16 or
    64 hot sources, so that most edges of these small graphs take the cold
    path (with the default 16,384 hot sources every source of a graph this
    small is hot)."""
    ctx = gbx.Context(0)
    t = ctx.tuning()
    t.binned_chunk_edges = 96
    t.binned_bin_edges = 64
    t.binned_bin_rows = 37
    t.binned_chunk_rows = 50
    t.binned_hot_sources = request.param
    ctx.set_tuning(t)
    yield ctx
    ctx.close()


# hot sources: the default (every source of these small graphs is hot, so
# only the hot pairs carry edges) and 16 (most edges through the scatter)
HOT = [16384, 16]


@pytest.mark.parametrize("hot", HOT)
@pytest.mark.parametrize("name", GRAPHS)
def test_binned_ref_matches_reference(gpu_ctx, golden, tuning, name, hot):
    tuning(binned_hot_sources=hot)
    z = golden(name)
    print(f"[gpu] binned REF {name}", flush=True)
    g = _graph(gpu_ctx, z)
    for iters, key in ((1, "pr_r1"), (20, "pr_r20")):
        ranks, rep = g.pagerank(iterations=iters, epsilon=0.0, method=BINNED)
        assert rep["executed"] == iters
        assert l1(ranks, z[key]) <= TOL_F32
    assert rep["error_total"] == pytest.approx(float(z["pr_err20"]), rel=1e-4)
    fwd, inv = _oracle_csr(z)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, O.row_ptr_from(inv), inv.dst, mode=0, iterations=20)
    err64 = l1(ranks, r64)
    print(f"[gpu] binned {name}: L1 vs reference {l1(ranks, z['pr_r20']):.3e}, vs FP64 {err64:.3e}")
    assert err64 <= TOL_F32


@pytest.mark.parametrize("name", ["rmat_s10", "rmat_s12", "fixture_v1000_e10000"])
def test_binned_ref_epsilon_stop(gpu_ctx, golden, name):
    z = golden(name)
    g = _graph(gpu_ctx, z)
    ranks, rep = g.pagerank(iterations=100, epsilon=1e-4, method=BINNED)
    assert abs(rep["iterations"] - int(z["pr_eps_iters"])) <= 1 and rep["converged"] == 1
    if rep["iterations"] == int(z["pr_eps_iters"]):
        assert l1(ranks, z["pr_eps_ranks"]) <= TOL_F32


@pytest.mark.parametrize("hot", HOT)
@pytest.mark.parametrize("name", ["rmat_s12", "fixture_v8192_e32768"])
def test_binned_forward_direction(gpu_ctx, golden, tuning, name, hot):
    tuning(binned_hot_sources=hot)
    z = golden(name)
    g = _graph(gpu_ctx, z)
    ranks, _ = g.pagerank(iterations=20, epsilon=0.0, direction=gbx.PULL_FWD, method=BINNED)
    assert l1(ranks, z["pr_fwd_r20"]) <= TOL_F32


@pytest.mark.parametrize("hot", HOT)
@pytest.mark.parametrize("name", ["rmat_s10", "rmat_s12", "fixture_v8192_e32768"])
def test_binned_stochastic_matches_fp64(gpu_ctx, golden, tuning, name, hot):
    tuning(binned_hot_sources=hot)
    z = golden(name)
    g = _graph(gpu_ctx, z)
    fwd, inv = _oracle_csr(z)
    rp = O.row_ptr_from(inv)
    ranks, rep = g.pagerank(iterations=100, epsilon=1e-6, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1,
                            method=BINNED)
    r64, it64, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, epsilon=1e-6)
    assert rep["converged"] == 1 and abs(rep["executed"] - it64) <= 1
    assert abs(rep["rank_sum"] - 1.0) < 1e-5
    assert l1(ranks, r64) <= TOL_F32
    # personalised teleport
    rng = np.random.default_rng(3)
    t = np.zeros(fwd.N)
    t[rng.choice(fwd.N, 40, replace=False)] = 1.0 / 40
    ranks, _ = g.pagerank(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, teleport=t,
                          method=BINNED)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=200, epsilon=1e-7, teleport=t)
    assert l1(ranks, r64) <= TOL_F32


def _edge_values(ctx, on):
    t = ctx.tuning()
    t.binned_edge_values = int(on)
    ctx.set_tuning(t)


@pytest.mark.parametrize("name", ["rmat_s12", "fixture_v8192_e32768", "fixture_v500_e500"])
def test_binned_small_cut_layout(small_cut_ctx, golden, name):
    """Thousands of segments, bins of <= 32 rows, chunks of <= 48 sources,
    rows split over several accumulate blocks.  Run layout: the cut changes
    which edges one FP32 run sum takes, so ranks agree with the default cut's
    (every source hot here) to rounding.  One V value per edge: exact integer
    sums make them bit-identical."""
    z = golden(name)
    print(f"[gpu] binned small cut {name}", flush=True)
    g = _graph(small_cut_ctx, z)
    ranks, _ = g.pagerank(iterations=20, epsilon=0.0, method=BINNED)
    assert l1(ranks, z["pr_r20"]) <= TOL_F32
    _edge_values(small_cut_ctx, True)
    ranks_e, _ = g.pagerank(iterations=20, epsilon=0.0, method=BINNED)
    assert l1(ranks_e, z["pr_r20"]) <= TOL_F32
    g2 = gbx.Context(0)
    try:
        g2g = _graph(g2, z)
        ref, _ = g2g.pagerank(iterations=20, epsilon=0.0, method=BINNED)
        _edge_values(g2, True)
        ref_e, _ = g2g.pagerank(iterations=20, epsilon=0.0, method=BINNED)
    finally:
        g2.close()
    assert l1(ranks, ref) <= 1e-6
    np.testing.assert_array_equal(ranks_e, ref_e)


@pytest.mark.parametrize("precision, tol", [(gbx.F32, 1e-6), (gbx.BF16, 2e-3)])
def test_binned_run_layout_matches_edge_values(gpu_ctx, golden, precision, tol):
    """The run layout (one V value per run of edges from one chunk into one
    row, summed in the scatter) against one V value per edge, same graph:
    REF and stochastic ranks agree to the rounding of the run sums."""
    z = golden("rmat_s12")
    g = _graph(gpu_ctx, z)
    out = {}
    for ev in (0, 1):
        _edge_values(gpu_ctx, ev)
        for mode in (gbx.PR_REF, gbx.PR_STOCHASTIC):
            out[ev, mode], _ = g.pagerank(iterations=20, epsilon=0.0, mode=mode, precision=precision,
                                          method=BINNED)
    _edge_values(gpu_ctx, 0)
    for mode in (gbx.PR_REF, gbx.PR_STOCHASTIC):
        e = l1(out[0, mode], out[1, mode])
        print(f"[gpu] run layout vs edge values, precision {precision} mode {mode}: L1 {e:.3e}")
        assert e <= tol


def test_binned_deterministic_and_graph_replay(gpu_ctx, golden):
    z = golden("rmat_s12")
    g = _graph(gpu_ctx, z)
    s = g.session(iterations=20, epsilon=0.0, stop=gbx.STOP_NONE, method=BINNED)
    s.reset()
    s.launch(20)
    r1, _ = s.read()
    s.reset()
    s.launch(20)
    r2, _ = s.read()
    np.testing.assert_array_equal(r1, r2)
    for rep in range(3):
        s.reset()
        s.launch_graph(5)
        s.launch_graph(15)
        r3, rp3 = s.read()
        assert rp3["executed"] == 20
        np.testing.assert_array_equal(r3, r1)


def test_binned_bf16(gpu_ctx, golden):
    z = golden("rmat_s12")
    g = _graph(gpu_ctx, z)
    fwd, inv = _oracle_csr(z)
    ranks, _ = g.pagerank(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1,
                          precision=gbx.BF16, method=BINNED)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, O.row_ptr_from(inv), inv.dst, mode=1, iterations=100,
                         epsilon=1e-7)
    e = l1(ranks, r64)
    print(f"[gpu] binned BF16 rmat_s12: L1 vs FP64 {e:.3e}")
    assert e <= 2e-3


def test_binned_auto_selection_and_errors(golden):
    z = golden("rmat_s10")
    ctx = gbx.Context(0)
    try:
        g = _graph(ctx, z)
        pull, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_PULL)
        binned, _ = g.pagerank(iterations=20, epsilon=0.0, method=BINNED)
        auto_small, _ = g.pagerank(iterations=20, epsilon=0.0)  # below the default threshold: pull
        np.testing.assert_array_equal(auto_small, pull)
        t = ctx.tuning()
        t.binned_min_vertices = 1
        ctx.set_tuning(t)
        auto, _ = g.pagerank(iterations=20, epsilon=0.0)
        np.testing.assert_array_equal(auto, binned)
        assert l1(binned, pull) <= 1e-6
        w = np.ones(len(z["edges_src"]), np.float32)
        gw = ctx.graph(z["edges_src"], z["edges_dst"], w, sort_mode=1, flags=gbx.GRAPH_WEIGHTED)
        # weighted stochastic runs on the binned method too (FP32); BF16 weighted on the pull method only
        rw_b, _ = gw.pagerank(iterations=5, mode=gbx.PR_STOCHASTIC, weighted=True, method=BINNED)
        rw_p, _ = gw.pagerank(iterations=5, mode=gbx.PR_STOCHASTIC, weighted=True, method=gbx.METHOD_PULL)
        rw_a, _ = gw.pagerank(iterations=5, mode=gbx.PR_STOCHASTIC, weighted=True)  # AUTO from 1 vertex: binned
        assert l1(rw_b, rw_p) <= 1e-6
        np.testing.assert_array_equal(rw_a, rw_b)
        with pytest.raises(gbx.GbxError, match="binned"):
            gw.pagerank(iterations=5, mode=gbx.PR_STOCHASTIC, weighted=True, precision=gbx.BF16, method=BINNED)
        gw.pagerank(iterations=5, mode=gbx.PR_STOCHASTIC, weighted=True, precision=gbx.BF16)  # AUTO: pull
    finally:
        ctx.close()


def test_binned_rmat_s20_matches_pull(gpu_ctx):
    """A full-size-shaped case: R-MAT s20 (1M vertices, ~16M edges) with the
    default cut, against the pull method and FP64."""
    g = gpu_ctx.rmat_graph(20, 16, 20261203 + 20)
    print(f"[gpu] binned R-MAT s20 N={g.N} E={g.E}", flush=True)
    pull, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_PULL)
    binned, _ = g.pagerank(iterations=20, epsilon=0.0, method=BINNED)
    inv = g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    fo = g.export()
    r64, _, _ = O.pr_f64(g.N, fo["out_degree"], rp, inv["dst"], mode=0, iterations=20)
    e_b, e_p = l1(binned, r64), l1(pull, r64)
    print(f"[gpu] s20 REF L1 vs FP64: binned {e_b:.3e}, pull {e_p:.3e}")
    assert e_b <= TOL_F32 and e_p <= TOL_F32


def test_binned_rmat_s24_matches_pull(gpu_ctx):
    """The benchmark's graph at its full size: R-MAT s24 (16.7M vertices,
    263M edges), seed 24, the default cut (the largest accumulate units, rows
    split over several units, units handed out largest first).  The binned
    iteration against the pull kernel after 5 iterations: FP32 within FP32
    rounding of the run sums, BF16 within its bar, and the same rank sum
    (below 1 under the reference's update, which drops dangling mass)."""
    g = gpu_ctx.rmat_graph(24, 16, 24)
    print(f"[gpu] binned R-MAT s24 N={g.N} E={g.E}", flush=True)
    pull, rp = g.pagerank(iterations=5, epsilon=0.0, method=gbx.METHOD_PULL)
    binned, rb = g.pagerank(iterations=5, epsilon=0.0, method=BINNED)
    bf16, r16 = g.pagerank(iterations=5, epsilon=0.0, method=BINNED, precision=gbx.BF16)
    e32, e16 = l1(binned, pull), l1(bf16, pull)
    print(f"[gpu] s24 binned vs pull: FP32 {e32:.3e}, BF16 {e16:.3e}", flush=True)
    assert e32 <= 1e-6
    assert e16 <= 2e-3
    for rep in (rp, rb, r16):
        assert rep["executed"] == 5
    assert abs(rb["rank_sum"] - rp["rank_sum"]) < 1e-6
    assert abs(r16["rank_sum"] - rp["rank_sum"]) < 2e-3
    g.free()
