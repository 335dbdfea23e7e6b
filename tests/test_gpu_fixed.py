"""The reference's fixed-point pull PageRank on the GPU
(pageRankPullFixedPoint{64,32,16,8}BitGraphCSR, pageRank.c:496-1030;
pushpull codes 2, 9, 10, 11, and 3 whose push form sums the same integers as
2): bit-exact ranks and identical stopping iterations against the reference's
own outputs, since the row sums are exact integer sums."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL]
MODES = ((gbx.PR_FIXED64, 2, 64), (gbx.PR_FIXED32, 9, 32), (gbx.PR_FIXED16, 10, 16), (gbx.PR_FIXED8, 11, 8))


def _graph(ctx, z):
    return ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)


@pytest.mark.parametrize("name", GRAPHS)
def test_fixed_point_bit_exact_vs_reference(gpu_ctx, golden, name):
    z, v = golden(name), golden(f"variants_{name}")
    print(f"[gpu] fixed point {name}", flush=True)
    g = _graph(gpu_ctx, z)
    for mode, code, bits in MODES:
        ranks, rep = g.pagerank(iterations=20, epsilon=0.0, mode=mode)
        np.testing.assert_array_equal(ranks, v[f"pp{code}_r20"], err_msg=f"{bits}-bit, 20 iterations")
        assert rep["error_total"] == pytest.approx(float(v[f"pp{code}_err20"]), rel=1e-6)
        ranks, rep = g.pagerank(iterations=100, epsilon=1e-4, mode=mode)
        assert rep["iterations"] == int(v[f"pp{code}_eps_iters"]), f"{bits}-bit stop"
        np.testing.assert_array_equal(ranks, v[f"pp{code}_eps_ranks"], err_msg=f"{bits}-bit, epsilon stop")
    # the push form of the 64-bit variant sums the same integers
    np.testing.assert_array_equal(v["pp3_r20"], v["pp2_r20"])


@pytest.mark.parametrize("name", ["rmat_s12", "fixture_v8192_e32768"])
def test_fixed_point_any_cut_is_bit_exact(golden, name):
    """Tiny chunks and bins (thousands of segments, rows split over several
    accumulate units, few hot sources): still the reference's bits."""
    z, v = golden(name), golden(f"variants_{name}")
    ctx = gbx.Context(0)
    try:
        t = ctx.tuning()
        t.binned_chunk_edges, t.binned_bin_edges, t.binned_bin_rows = 96, 64, 37
        t.binned_chunk_rows, t.binned_hot_sources = 48, 16
        ctx.set_tuning(t)
        g = _graph(ctx, z)
        for mode, code, bits in MODES:
            ranks, _ = g.pagerank(iterations=20, epsilon=0.0, mode=mode)
            np.testing.assert_array_equal(ranks, v[f"pp{code}_r20"], err_msg=f"{bits}-bit")
    finally:
        ctx.close()


def test_fixed_point_rmat_s18_vs_oracle(gpu_ctx):
    g = gpu_ctx.rmat_graph(18, 16, 20261203 + 18)
    print(f"[gpu] fixed point R-MAT s18 N={g.N} E={g.E}", flush=True)
    f, inv = g.export(), g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    for mode, _, bits in MODES:
        ranks, rep = g.pagerank(iterations=100, epsilon=1e-4, mode=mode)
        r64, it, _ = O.pr_pull_fixed(g.N, f["out_degree"], rp, inv["dst"], bits, 100, 1e-4)
        assert rep["iterations"] == it
        np.testing.assert_array_equal(ranks, r64, err_msg=f"{bits}-bit")


def test_fixed_point_graph_replay_and_errors(gpu_ctx, golden):
    z, v = golden("rmat_s10"), golden("variants_rmat_s10")
    g = _graph(gpu_ctx, z)
    s = g.session(iterations=20, epsilon=0.0, stop=gbx.STOP_NONE, mode=gbx.PR_FIXED16)
    for _ in range(3):
        s.reset()
        s.launch_graph(7)
        s.launch_graph(13)
        r, rep = s.read()
        assert rep["executed"] == 20
        np.testing.assert_array_equal(r, v["pp10_r20"])
    s.free()
    with pytest.raises(gbx.GbxError, match="binned"):
        g.pagerank(iterations=5, mode=gbx.PR_FIXED32, method=gbx.METHOD_PULL)
    with pytest.raises(gbx.GbxError, match="teleport"):
        g.pagerank(iterations=5, mode=gbx.PR_FIXED32, teleport=np.full(g.N, 1.0 / g.N))
