"""The reference's quantized pull PageRank on the GPU
(pageRankPullQuant{32,16,8}BitGraphCSR, pageRank.c:1159-1577; pushpull codes
4, 12, 13): bit-exact ranks and identical stopping iterations against the
reference's own outputs.  The contributions' exact min and max come out of
the previous iteration's reduction, the scatter and the hot copy quantize
them with the reference's scale and rounding, and the row sums are exact
integer sums.  ring_rmat_s10 (no dangling vertex) makes the 32-bit
quantization wrap past 2^32 and the 16/8-bit one saturate; ring_64 has every
contribution equal (scale 1)."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL]
RINGS = ["ring_rmat_s10", "ring_64"]
MODES = ((gbx.PR_QUANT32, 4, 32), (gbx.PR_QUANT16, 12, 16), (gbx.PR_QUANT8, 13, 8))


def _graph(ctx, golden, name):
    if name in RINGS:
        v = golden(f"variants_{name}")
        return ctx.graph(v["edges_src"], v["edges_dst"], sort_mode=1)
    z = golden(name)
    return ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)


@pytest.mark.parametrize("name", GRAPHS + RINGS)
def test_quantized_bit_exact_vs_reference(gpu_ctx, golden, name):
    v = golden(f"variants_{name}")
    print(f"[gpu] quantized {name}", flush=True)
    g = _graph(gpu_ctx, golden, name)
    for mode, code, bits in MODES:
        ranks, rep = g.pagerank(iterations=20, epsilon=0.0, mode=mode)
        np.testing.assert_array_equal(ranks, v[f"pp{code}_r20"], err_msg=f"{bits}-bit, 20 iterations")
        assert rep["error_total"] == pytest.approx(float(v[f"pp{code}_err20"]), rel=1e-6, abs=1e-300)
        ranks, rep = g.pagerank(iterations=100, epsilon=1e-4, mode=mode)
        assert rep["iterations"] == int(v[f"pp{code}_eps_iters"]), f"{bits}-bit stop"
        np.testing.assert_array_equal(ranks, v[f"pp{code}_eps_ranks"], err_msg=f"{bits}-bit, epsilon stop")


@pytest.mark.parametrize("name", ["rmat_s12", "ring_rmat_s10"])
def test_quantized_any_cut_is_bit_exact(golden, name):
    """Tiny chunks and bins (thousands of segments, rows split over several
    accumulate units, few hot sources): still the reference's bits."""
    v = golden(f"variants_{name}")
    ctx = gbx.Context(0)
    try:
        t = ctx.tuning()
        t.binned_chunk_edges, t.binned_bin_edges, t.binned_bin_rows = 96, 64, 37
        t.binned_chunk_rows, t.binned_hot_sources = 48, 16
        ctx.set_tuning(t)
        g = _graph(ctx, golden, name)
        for mode, code, bits in MODES:
            ranks, _ = g.pagerank(iterations=20, epsilon=0.0, mode=mode)
            np.testing.assert_array_equal(ranks, v[f"pp{code}_r20"], err_msg=f"{bits}-bit")
    finally:
        ctx.close()


def test_quantized_rmat_s18_vs_oracle(gpu_ctx):
    g = gpu_ctx.rmat_graph(18, 16, 20261203 + 18)
    print(f"[gpu] quantized R-MAT s18 N={g.N} E={g.E}", flush=True)
    f, inv = g.export(), g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    for mode, _, bits in MODES:
        ranks, rep = g.pagerank(iterations=100, epsilon=1e-4, mode=mode)
        ro, it, _ = O.pr_pull_quant(g.N, f["out_degree"], rp, inv["dst"], bits, 100, 1e-4)
        assert rep["iterations"] == it
        np.testing.assert_array_equal(ranks, ro, err_msg=f"{bits}-bit")


def test_quantized_graph_replay_and_errors(gpu_ctx, golden):
    v = golden("variants_rmat_s10")
    g = _graph(gpu_ctx, golden, "rmat_s10")
    s = g.session(iterations=20, epsilon=0.0, stop=gbx.STOP_NONE, mode=gbx.PR_QUANT16)
    for _ in range(3):
        s.reset()
        s.launch_graph(7)
        s.launch_graph(13)
        r, rep = s.read()
        assert rep["executed"] == 20
        np.testing.assert_array_equal(r, v["pp12_r20"])
    s.free()
    with pytest.raises(gbx.GbxError, match="binned"):
        g.pagerank(iterations=5, mode=gbx.PR_QUANT8, method=gbx.METHOD_PULL)
    with pytest.raises(gbx.GbxError, match="teleport"):
        g.pagerank(iterations=5, mode=gbx.PR_QUANT32, teleport=np.full(g.N, 1.0 / g.N))
