"""Batches of independent graphs in one launch per iteration, and iterations
replayed as HIP graphs."""
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


def test_batch_matches_each_graph_alone(gpu_ctx):
    members = _members(gpu_ctx, [6, 10, 8, 12, 7, 11, 9, 13], seed=1)
    b = gpu_ctx.graph_batch(members)
    off = b.batch_offsets()
    assert len(off) == len(members) + 1 and off[-1] == b.N == sum(m.N for m in members)
    s = b.session(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
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


def test_batch_of_256_graphs_log_uniform(gpu_ctx):
    """Config shape: 256 graphs, N log-uniform, edge factor 8..32 (sizes capped for the test)."""
    rng = np.random.default_rng(256)
    scales = np.clip(np.round(rng.uniform(10, 15, 256)), 10, 15).astype(int)
    members = _members(gpu_ctx, scales, seed=2)
    b = gpu_ctx.graph_batch(members)
    s = b.session(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
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


@pytest.mark.parametrize("form", ["single_ref", "single_bf16", "batched_vectors", "graph_batch"])
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
        s = g.session(iterations=30, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE)
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
