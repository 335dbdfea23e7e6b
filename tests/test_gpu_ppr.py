"""Batched personalised PageRank (k vectors per launch) against the FP64
oracle run one vector at a time."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _seed_sets(rng, N, k, lo=1, hi=1000):
    return [rng.choice(N, int(rng.integers(lo, min(hi, N) + 1)), replace=False) for _ in range(k)]


def _oracle(z):
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    return fwd, inv, O.row_ptr_from(inv)


@pytest.mark.parametrize("k", [1, 3, 4, 8, 32])
def test_batched_ppr_matches_single_vector_fp64(gpu_ctx, golden, k):
    z = golden("rmat_s12")
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    fwd, inv, rp = _oracle(z)
    rng = np.random.default_rng(100 + k)
    sets = _seed_sets(rng, fwd.N, k, 1, 300)
    s = g.ppr_session(k, sets, iterations=200, epsilon=1e-7)
    s.reset()
    s.launch(200)
    ranks, reps = s.read()
    assert ranks.shape == (fwd.N, k)
    for q in range(k):
        t = np.zeros(fwd.N)
        t[sets[q]] = 1.0 / len(sets[q])
        r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=300, epsilon=1e-12, teleport=t)
        e = l1(ranks[:, q], r64)
        assert e <= 1e-5, (q, e)
        assert reps[q]["converged"] == 1
        assert abs(reps[q]["rank_sum"] - 1.0) < 1e-5


def test_batched_uniform_equals_pagerank(gpu_ctx, golden):
    z = golden("fixture_v8192_e32768")
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    single, _ = g.pagerank(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
    s = g.ppr_session(5, None, iterations=100, epsilon=1e-7)
    s.reset()
    s.launch(100)
    ranks, reps = s.read()
    for q in range(5):
        assert l1(ranks[:, q], single) <= 2e-6
        np.testing.assert_array_equal(ranks[:, q], ranks[:, 0])


@pytest.mark.parametrize("vpl", [1, 2, 4])
def test_batched_hub_chunks_and_full_width(gpu_ctx, tuning, vpl):
    """Scale-16 graph with small hub threshold/chunks so every path is used;
    64 vectors, 1, 2 or 4 of them a lane (teams of 64, 32, 16 lanes)."""
    tuning(ppr_hub=128, ppr_chunk=256, ppr_vectors_per_lane=vpl)
    g = gpu_ctx.rmat_graph(16, 16, 77)
    e = g.export(False), g.export(True)
    fwd = O.OracleCSR(g.N, g.E, e[0]["out_degree"], e[0]["in_degree"], e[0]["edges_idx"], e[0]["src"], e[0]["dst"], None)
    inv = O.OracleCSR(g.N, g.E, e[1]["out_degree"], e[1]["in_degree"], e[1]["edges_idx"], e[1]["src"], e[1]["dst"], None)
    rp = O.row_ptr_from(inv)
    rng = np.random.default_rng(64)
    sets = _seed_sets(rng, g.N, 64)
    s = g.ppr_session(64, sets, iterations=60, epsilon=0.0, stop=gbx.STOP_NONE)
    info = s.ppr_info()
    assert info["vector_width"] == 64 // vpl and info["hub_chunks"] > info["hub_rows"] > 0
    s.reset()
    s.launch(60)
    ranks, reps = s.read()
    assert reps[0]["executed"] == 60
    for q in (0, 17, 63):
        t = np.zeros(g.N)
        t[sets[q]] = 1.0 / len(sets[q])
        r64, _, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=60, epsilon=0.0, teleport=t)
        assert l1(ranks[:, q], r64) <= 1e-5


def test_batched_rejects_bad_input(gpu_ctx, golden):
    z = golden("rmat_s8")
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    with pytest.raises(gbx.GbxError):
        g.ppr_session(65, None)
    with pytest.raises(gbx.GbxError):
        g.ppr_session(2, [np.array([1, 2]), np.array([], np.uint32)])
    with pytest.raises(gbx.GbxError):
        g.ppr_session(1, [np.array([10 ** 6])])
