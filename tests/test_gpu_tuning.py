"""Tuning interface, autotuner and batched vector widths: every setting gives
results within the same tolerances."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture()
def fresh_ctx():
    ctx = gbx.Context(0)  # own context: tuning changes stay local
    yield ctx
    ctx.close()


def test_tuning_roundtrip_and_validation(fresh_ctx):
    t = fresh_ctx.tuning()
    d = t.as_dict()
    assert d["bin_thresholds"] == [0, 256, 128, 64, 32, 16, 8] and d["hub_chunk"] == 0 and d["lds_bytes"] == 159744
    t.hub_chunk = 16384
    t.ppr_vectors_per_lane = 2
    fresh_ctx.set_tuning(t)
    assert fresh_ctx.tuning().hub_chunk == 16384
    bad = fresh_ctx.tuning()
    bad.bin_thresholds[3] = 100000  # not descending
    with pytest.raises(gbx.GbxError):
        fresh_ctx.set_tuning(bad)
    bad = fresh_ctx.tuning()
    bad.ppr_vectors_per_lane = 3
    with pytest.raises(gbx.GbxError):
        fresh_ctx.set_tuning(bad)


@pytest.mark.parametrize("variant", ["small_chunks", "no_lds", "two_blocks", "wide_lanes"])
def test_plan_rebuilds_and_results_hold_under_tuning(fresh_ctx, golden, variant):
    z = golden("rmat_s12")
    g = fresh_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    r0, _ = g.pagerank(iterations=20, epsilon=0.0)
    t = fresh_ctx.tuning()
    if variant == "small_chunks":
        t.bin_thresholds[0] = 256
        t.hub_chunk = 256
    elif variant == "no_lds":
        t.lds_bytes = 0
    elif variant == "two_blocks":
        t.blocks_per_cu = 2
        t.lds_bytes = 65536
    else:
        for i, v in enumerate([512, 256, 128, 64, 32, 16]):
            t.bin_thresholds[i + 1] = v
    fresh_ctx.set_tuning(t)
    r1, _ = g.pagerank(iterations=20, epsilon=0.0)
    assert l1(r1, z["pr_r20"]) <= 1e-5
    if variant == "no_lds":
        np.testing.assert_array_equal(r1, r0)


@pytest.mark.parametrize("vpl", [1, 2, 4])
@pytest.mark.parametrize("k", [4, 16, 32])
def test_batched_vectors_per_lane(fresh_ctx, golden, vpl, k):
    z = golden("rmat_s12")
    t = fresh_ctx.tuning()
    t.ppr_vectors_per_lane = vpl
    t.ppr_hub = 128
    t.ppr_chunk = 512
    fresh_ctx.set_tuning(t)
    g = fresh_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    rp = O.row_ptr_from(inv)
    rng = np.random.default_rng(k * 10 + vpl)
    sets = [rng.choice(g.N, int(rng.integers(1, 200)), replace=False) for _ in range(k)]
    s = g.ppr_session(k, sets, iterations=100, epsilon=0.0, stop=gbx.STOP_NONE)
    assert s.ppr_info()["vector_width"] == max(1, (1 << (k - 1).bit_length()) // vpl)
    s.reset()
    s.launch(100)
    ranks, reps = s.read()
    for q in (0, k - 1):
        tv = np.zeros(g.N)
        tv[sets[q]] = 1.0 / len(sets[q])
        r64, _, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, epsilon=0.0, teleport=tv)
        assert l1(ranks[:, q], r64) <= 1e-5


def test_autotune_single_and_batched(fresh_ctx):
    g = fresh_ctx.rmat_graph(16, 16, 9)
    ref, _ = g.pagerank(iterations=30, epsilon=0.0)
    best, ms = fresh_ctx.autotune(g, 0, 3, iterations=10, epsilon=0.0)
    assert ms > 0 and fresh_ctx.tuning().as_dict() == best.as_dict()
    r, _ = g.pagerank(iterations=30, epsilon=0.0)
    assert l1(r, ref) <= 1e-5
    bestb, msb = fresh_ctx.autotune(g, 8, 3, iterations=10, epsilon=0.0, mode=gbx.PR_STOCHASTIC)
    assert msb > 0 and bestb.ppr_vectors_per_lane in (1, 2, 4)
    print("[gpu] autotuned:", best.as_dict(), ms, bestb.as_dict(), msb)


def test_autotune_binned_layout_cut(fresh_ctx):
    """With the binned iteration the autotuner searches its layout cut and hot
    sources (not the pull kernel's fields), and the ranks stay within FP32
    rounding of the default cut's (the run sums follow the cut)."""
    g = fresh_ctx.rmat_graph(17, 16, 12)
    start = fresh_ctx.tuning().as_dict()
    ref, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_BINNED)
    best, ms = fresh_ctx.autotune(g, 0, 3, iterations=10, epsilon=0.0, method=gbx.METHOD_BINNED)
    b = best.as_dict()
    assert ms > 0 and fresh_ctx.tuning().as_dict() == b
    for k in ("bin_thresholds", "hub_chunk", "lds_bytes", "blocks_per_cu"):
        assert b[k] == start[k], k
    r, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_BINNED)
    assert l1(r, ref) <= 1e-6
    print("[gpu] autotuned binned:", {k: b[k] for k in b if k.startswith("binned")}, ms)


def test_autotune_graph_batch(fresh_ctx):
    """A batch of graphs tunes its tile cut and grid; every member's ranks
    stay within 1e-6 of those under the starting setting."""
    members = [fresh_ctx.rmat_graph(s, 8, 70 + s) for s in (10, 12, 13, 11)]
    b = fresh_ctx.graph_batch(members)
    kw = dict(iterations=15, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE)
    s = b.session(**kw)
    s.reset()
    s.launch(15)
    ref, _ = s.read()
    s.free()
    start = fresh_ctx.tuning().as_dict()
    best, ms = fresh_ctx.autotune(b, 0, 3, **kw)
    d = best.as_dict()
    assert ms > 0 and fresh_ctx.tuning().as_dict() == d
    for k in ("lds_bytes", "blocks_per_cu", "binned_chunk_edges", "ppr_vectors_per_lane"):
        assert d[k] == start[k], k
    s = b.session(**kw)
    s.reset()
    s.launch(15)
    r, reps = s.read()
    s.free()
    off = b.batch_offsets()
    for k in range(len(members)):
        assert l1(r[off[k]:off[k + 1]], ref[off[k]:off[k + 1]]) <= 1e-6, k
        assert abs(reps[k]["rank_sum"] - 1.0) < 1e-5
    b.free()
    for m in members:
        m.free()
