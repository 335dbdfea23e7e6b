"""GPU CSR build: bit-exact against the reference's arrays (golden vectors)."""
import hashlib

import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL]


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_csr_build_bit_exact(gpu_ctx, golden, name, mode):
    z = golden(name)
    tag = "m1" if mode in (1, 3) else "m2"
    N = int(z[f"{tag}_N"])
    print(f"[gpu] csr_build {name} mode={mode} N={N} E={z['edges_src'].size}", flush=True)
    out = gpu_ctx.csr_build(z["edges_src"], z["edges_dst"], None, N, mode)
    np.testing.assert_array_equal(out["src"], z[f"{tag}_fwd_src"])
    np.testing.assert_array_equal(out["dst"], z[f"{tag}_fwd_dst"])
    np.testing.assert_array_equal(out["out_degree"], z[f"{tag}_fwd_out_degree"])
    np.testing.assert_array_equal(out["edges_idx"], z[f"{tag}_fwd_edges_idx"])


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("tag,mode,flags", [("m1", 1, 0), ("m2", 2, 0), ("k", 1, gbx.GRAPH_DEDUPE)])
def test_graph_forward_and_transpose_bit_exact(gpu_ctx, golden, name, tag, mode, flags):
    z = golden(name)
    print(f"[gpu] graph {name} {tag}", flush=True)
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=mode, flags=flags)
    assert (g.N, g.E) == (int(z[f"{tag}_N"]), int(z[f"{tag}_E"]))
    for inv, side in ((False, "fwd"), (True, "inv")):
        e = g.export(inverse=inv)
        for k in ("out_degree", "in_degree", "edges_idx", "src", "dst"):
            np.testing.assert_array_equal(e[k], z[f"{tag}_{side}_{k}"], err_msg=f"{side}.{k}")
    g.free()


@pytest.mark.parametrize("fx", FIXTURES)
def test_graph_matches_shipped_text_dumps(gpu_ctx, golden, fx):
    t = golden(f"textdump_{fx}")
    z = golden(f"fixture_{fx}")
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=3)
    e = g.export()
    assert g.N == int(t["N"]) and g.E == int(t["E"])
    for k in ("out_degree", "edges_idx", "src", "dst"):
        np.testing.assert_array_equal(e[k], t[k])


@pytest.mark.parametrize("s", RMAT_SMALL)
@pytest.mark.parametrize("tag,mode", [("wm1", 1), ("wm2", 2)])
def test_weights_follow_the_sort(gpu_ctx, golden, s, tag, mode):
    z = golden(f"rmat_s{s}")
    g = gpu_ctx.graph(z["wedges_src"], z["wedges_dst"], z["wedges_w"], sort_mode=mode, flags=gbx.GRAPH_WEIGHTED)
    for inv, side in ((False, "fwd"), (True, "inv")):
        e = g.export(inverse=inv)
        np.testing.assert_array_equal(e["w"], z[f"{tag}_{side}_w"])
        np.testing.assert_array_equal(e["dst"], z[f"{tag}_{side}_dst"])


def test_stable_sort_keeps_input_order_within_rows(gpu_ctx):
    rng = np.random.default_rng(5)
    n, E = 3000, 200000
    src = rng.integers(0, n, E).astype(np.uint32)
    dst = rng.integers(0, n, E).astype(np.uint32)
    w = np.arange(E, dtype=np.float32)
    out = gpu_ctx.csr_build(src, dst, w, n, 2, move_weights=True)
    np.testing.assert_array_equal(out["w"], w[np.argsort(src, kind="stable")])
    out3 = gpu_ctx.csr_build(src, dst, w, n, 3, move_weights=True)
    np.testing.assert_array_equal(out3["w"], w[np.lexsort((dst, src))])


def test_symmetric_flag(gpu_ctx, golden):
    z = golden("fixture_v500_e500")
    s, d = z["edges_src"], z["edges_dst"]
    g = gpu_ctx.graph(s, d, sort_mode=1, flags=gbx.GRAPH_SYMMETRIC)
    fwd, _ = O.build_graph(np.concatenate([s, d]), np.concatenate([d, s]), None, g.N, mode=1)
    e = g.export()
    assert g.E == 2 * s.size
    np.testing.assert_array_equal(e["src"], fwd.src)
    np.testing.assert_array_equal(e["dst"], fwd.dst)


def test_edge_cases(gpu_ctx):
    # empty edge list: N = 1 (max id 0 + 1), all zeros
    g = gpu_ctx.graph(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert (g.N, g.E) == (1, 0)
    # only self loops with dedupe -> no edges left, N kept
    g2 = gpu_ctx.graph(np.array([3, 3, 5], np.uint32), np.array([3, 3, 5], np.uint32), flags=gbx.GRAPH_DEDUPE)
    assert (g2.N, g2.E) == (6, 0)
    e = g2.export(inverse=True)
    assert not e["out_degree"].any()
    # ids beyond 30 bits and beyond num_vertices are errors
    with pytest.raises(gbx.GbxError) as ei:
        gpu_ctx.graph(np.array([1 << 30], np.uint32), np.array([0], np.uint32))
    assert ei.value.status == -4
    with pytest.raises(gbx.GbxError):
        gpu_ctx.csr_build(np.array([0, 9], np.uint32), np.array([1, 1], np.uint32), None, 5, 1)
    with pytest.raises(gbx.GbxError):
        gpu_ctx.csr_build(np.array([0], np.uint32), np.array([1], np.uint32), None, 5, 7)
    # single vertex with a self loop
    g3 = gpu_ctx.graph(np.array([0], np.uint32), np.array([0], np.uint32))
    assert (g3.N, g3.E) == (1, 1)
    ranks, rep = g3.pagerank(iterations=20, epsilon=0.0)
    assert np.isfinite(ranks).all()


@pytest.mark.parametrize("scale", [8, 10, 12, 16])
def test_gpu_rmat_generator_matches_oracle(gpu_ctx, scale):
    seed = 20261203 + scale
    s, d, _ = gpu_ctx.rmat_edges(scale, 16, seed)
    os_, od, N = O.rmat_graph(scale, 16, seed)
    np.testing.assert_array_equal(s, os_)
    np.testing.assert_array_equal(d, od)


def test_gpu_rmat_graph_matches_golden_s16(gpu_ctx, golden):
    z = golden("rmat_s16")
    g = gpu_ctx.rmat_graph(16, 16, 20261203 + 16)
    assert (g.N, g.E) == (int(z["m1_N"]), int(z["m1_E"]))
    for inv, side in ((False, "fwd"), (True, "inv")):
        e = g.export(inverse=inv)
        for k in ("out_degree", "edges_idx", "src", "dst"):
            assert hashlib.sha256(e[k].tobytes()).hexdigest() == str(z[f"m1_{side}_{k}_sha"]), f"{side}.{k}"


def test_csr_invariants_at_scale_22(gpu_ctx):
    """Size-independent properties at the 4M-vertex workload size."""
    print("[gpu] rmat s22 graph build", flush=True)
    g = gpu_ctx.rmat_graph(22, 16, 22)
    assert g.E > 60_000_000
    for inv in (False, True):
        e = g.export(inverse=inv)
        od, idx, src, dst = e["out_degree"], e["edges_idx"], e["src"], e["dst"]
        assert int(od.sum(dtype=np.uint64)) == g.E
        assert np.all(np.diff(src.astype(np.int64)) >= 0)             # sorted by row
        key = src.astype(np.uint64) << np.uint64(32) | dst.astype(np.uint64)
        assert np.all(np.diff(key) > 0)                                # (src,dst) strictly increasing: deduped
        nz = od > 0
        start = np.zeros(g.N, np.int64)
        start[1:] = np.cumsum(od.astype(np.int64))[:-1]
        np.testing.assert_array_equal(idx[nz], start[nz])
        assert np.all(idx[~nz] == 0)
        np.testing.assert_array_equal(np.bincount(src, minlength=g.N), od)
    g.free()
