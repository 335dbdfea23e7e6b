"""Batches of independent graphs in one launch per iteration (graph-pure pull
tiles, or the binned iteration with bins cut at every graph's first row and
each unit updated with its graph's coefficients), and iterations replayed as
HIP graphs."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_of(g):
    f, i = g.export(False), g.export(True)
    fwd = O.OracleCSR(g.N, g.E, f["out_degree"], f["in_degree"], f["edges_idx"], f["src"], f["dst"], None)
    inv = O.OracleCSR(g.N, g.E, i["out_degree"], i["in_degree"], i["edges_idx"], i["src"], i["dst"], None)
    return fwd, inv, O.row_ptr_from(inv)


def _members(ctx, sizes, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k, s in enumerate(sizes):
        out.append(ctx.rmat_graph(int(s), int(rng.integers(8, 33)), 1000 + k))
    return out


def _tiny(ctx):
    """A 2-cycle, a star whose leaves have no out-edges (dangling mass) and a
    single self loop: graphs of 2, 9 and 1 vertices."""
    return [ctx.graph(np.array([0, 1], np.uint32), np.array([1, 0], np.uint32)),
            ctx.graph(np.zeros(8, np.uint32), np.arange(1, 9, dtype=np.uint32)),
            ctx.graph(np.array([0], np.uint32), np.array([0], np.uint32))]


@pytest.mark.parametrize("method", [gbx.METHOD_PULL, gbx.METHOD_BINNED])
def test_batch_matches_each_graph_alone(gpu_ctx, method):
    members = _members(gpu_ctx, [6, 10, 8, 12, 7, 11, 9, 13], seed=1) + _tiny(gpu_ctx)
    b = gpu_ctx.graph_batch(members)
    off = b.batch_offsets()
    assert len(off) == len(members) + 1 and off[-1] == b.N == sum(m.N for m in members)
    s = b.session(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, method=method)
    s.reset()
    s.launch(200)
    ranks, reps = s.read()
    assert len(reps) == len(members)
    for k, m in enumerate(members):
        rk = ranks[off[k]:off[k + 1]]
        alone, rep1 = m.pagerank(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
        fwd, inv, rp = _oracle_of(m)
        r64, it64, _ = O.pr_f64(m.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=200, epsilon=1e-7)
        assert reps[k]["converged"] == 1
        assert abs(reps[k]["executed"] - rep1["executed"]) <= 1
        assert abs(reps[k]["rank_sum"] - 1.0) < 1e-5
        assert l1(rk, r64) <= 1e-5, k
        assert l1(rk, alone) <= 2e-6, k


@pytest.mark.parametrize("method", [gbx.METHOD_PULL, gbx.METHOD_BINNED])
def test_batch_of_256_graphs_log_uniform(gpu_ctx, method):
    """Config shape: 256 graphs, N log-uniform, edge factor 8..32 (sizes capped for the test)."""
    rng = np.random.default_rng(256)
    scales = np.clip(np.round(rng.uniform(10, 15, 256)), 10, 15).astype(int)
    members = _members(gpu_ctx, scales, seed=2)
    b = gpu_ctx.graph_batch(members)
    s = b.session(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, method=method)
    s.reset()
    s.launch_graph(100)
    ranks, reps = s.read()
    off = b.batch_offsets()
    sums = np.array([ranks[off[k]:off[k + 1]].astype(np.float64).sum() for k in range(256)])
    assert np.abs(sums - 1.0).max() < 1e-5
    assert all(r["converged"] for r in reps)
    for k in (0, 99, 255):
        m = members[k]
        fwd, inv, rp = _oracle_of(m)
        r64, _, _ = O.pr_f64(m.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, epsilon=1e-7)
        assert l1(ranks[off[k]:off[k + 1]], r64) <= 1e-5


@pytest.mark.parametrize("form", ["single_ref", "single_bf16", "batched_vectors", "graph_batch",
                                  "graph_batch_binned"])
def test_graph_replay_is_bit_identical_to_eager(gpu_ctx, golden, form):
    z = golden("rmat_s12")
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    if form == "single_ref":
        s = g.session(iterations=30, epsilon=0.0, stop=gbx.STOP_NONE)
    elif form == "single_bf16":
        s = g.session(iterations=30, epsilon=0.0, stop=gbx.STOP_NONE, precision=gbx.BF16, mode=gbx.PR_STOCHASTIC)
    elif form == "batched_vectors":
        rng = np.random.default_rng(5)
        s = g.ppr_session(8, [rng.choice(g.N, 20, replace=False) for _ in range(8)], iterations=30, epsilon=0.0,
                          stop=gbx.STOP_NONE)
    else:
        g = gpu_ctx.graph_batch(_members(gpu_ctx, [8, 11, 9], seed=3) + [g])
        method = gbx.METHOD_BINNED if form == "graph_batch_binned" else gbx.METHOD_AUTO
        s = g.session(iterations=30, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE, method=method)
    s.reset()
    s.launch(10)
    eager, _ = s.read()
    for replay in range(3):
        s.reset()
        s.launch_graph(10)  # captured on the first replay, relaunched afterwards
        r, _ = s.read()
        np.testing.assert_array_equal(r, eager, err_msg=f"replay {replay}")
        # eager work and synchronisation between replays
        s.reset()
        s.launch(3)
        gpu_ctx.sync()
    # two captured halves (parity 0 then parity 1) equal ten eager iterations
    s.reset()
    s.launch(10)
    eager10, _ = s.read()
    s.reset()
    s.launch_graph(5)
    s.launch_graph(5)
    r, _ = s.read()
    np.testing.assert_array_equal(r, eager10)


def test_binned_batch_freezes_converged_graphs(gpu_ctx):
    """Each graph stops at its own L1 rule: after its stop its ranks stay as
    they were, its iteration count is its own, and the others run on; the
    stopping iterations agree with the pull batch's within one."""
    members = _members(gpu_ctx, [12, 7, 13, 9], seed=9) + _tiny(gpu_ctx)
    b = gpu_ctx.graph_batch(members)
    off = b.batch_offsets()
    out = {}
    for method in (gbx.METHOD_PULL, gbx.METHOD_BINNED):
        s = b.session(iterations=300, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, method=method)
        s.reset()
        s.launch(300)
        r300, reps = s.read()
        s.reset()
        s.launch(max(rp["executed"] for rp in reps) + 5)
        rmore, reps2 = s.read()
        np.testing.assert_array_equal(r300, rmore)  # converged graphs do not move
        assert [rp["executed"] for rp in reps] == [rp["executed"] for rp in reps2]
        out[method] = (r300, reps)
        s.free()
    (rp_, reps_p), (rb, reps_b) = out[gbx.METHOD_PULL], out[gbx.METHOD_BINNED]
    for k in range(len(members)):
        assert reps_b[k]["converged"] == 1 and abs(reps_b[k]["executed"] - reps_p[k]["executed"]) <= 1, k
        assert l1(rb[off[k]:off[k + 1]], rp_[off[k]:off[k + 1]]) <= 1e-6, k
