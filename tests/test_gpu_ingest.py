"""readEdgeListsbin (edgeList.c:450-706) on the GPU: the raw records of a
.bin / .wbin file to a device edge list, then the graph build from it.

* the reference's fixture files (their records in file order, tests/golden):
  the device edge list equals the host reader's arrays (the adapter's
  readEdgeListsbin, a separate restatement of the reference's loop) and the
  graphs built from it equal the reference's own CSR goldens (sort modes 1
  and 2, the -k dedupe);
* inverse, symmetric and both, from a file and from host memory, weighted
  records with NaN, negative, -0.0 and denormal weights (the largest weight
  as the reader's running `w > max` from 0.0f), a partial last record, an
  empty file;
* R-MAT s22 (65M edges, 0.5 GB of records, 0.8 GB weighted: the 64 MiB
  staging pieces) against the host reader, and its graph against the graph
  built from the host arrays;
* errors: a missing file, ids of 2^30 and beyond (the column mask)."""
import ctypes as C

import numpy as np
import pytest

import graphblox_amd as gbx
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIXTURES = ["v500_e500", "v1024_e4096", "v8192_e32768", "v1000_e10000", "v529_e3500"]


def _host_read(path, inverse=0, symmetric=0, weighted=0):
    """The adapter's readEdgeListsbin (weighted layout, so weights are read)."""
    import hostapi

    L = hostapi.load(True)
    e = L.readEdgeListsbin(path.encode(), inverse, symmetric, weighted).contents
    E = e.num_edges
    out = (np.ctypeslib.as_array(e.edges_array_src, (E,)).copy() if E else np.zeros(0, np.uint32),
           np.ctypeslib.as_array(e.edges_array_dest, (E,)).copy() if E else np.zeros(0, np.uint32),
           np.ctypeslib.as_array(e.edges_array_weight, (E,)).copy() if E else np.zeros(0, np.float32),
           e.num_vertices, e.max_weight)
    L.freeEdgeList(C.pointer(e))
    return out


def _same_edges(ed, host):
    s, d, w, N, mw = host
    gs, gd, gw = ed.export()
    np.testing.assert_array_equal(gs, s)
    np.testing.assert_array_equal(gd, d)
    np.testing.assert_array_equal(gw.view(np.uint32), w.view(np.uint32))  # bit patterns, NaN included
    assert ed.N == N and ed.E == s.size
    assert np.float32(ed.max_weight).view(np.uint32) == np.float32(mw).view(np.uint32)


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_files_match_reader_and_reference_csr(gpu_ctx, golden, tmp_path, name):
    z = golden("fixture_" + name)
    p = str(tmp_path / "graph.bin")
    O.write_bin(p, z["edges_src"], z["edges_dst"])
    ed = gpu_ctx.edges_read_file(p)
    _same_edges(ed, _host_read(p))
    assert ed.N == int(z["m1_N"])
    for mode, key, flags in ((gbx.SORT_RADIX_SRC_DST, "m1", 0), (gbx.SORT_COUNT_SRC, "m2", 0),
                             (gbx.SORT_RADIX_SRC_DST, "k", gbx.GRAPH_DEDUPE)):
        e2 = gpu_ctx.edges_read_file(p)
        g = e2.graph(sort_mode=mode, flags=flags)
        assert e2.E == 0  # consumed
        e2.free()
        for side, inv in (("fwd", False), ("inv", True)):
            x = g.export(inverse=inv)
            for k in ("out_degree", "in_degree", "edges_idx", "src", "dst"):
                np.testing.assert_array_equal(x[k], z[f"{key}_{side}_{k}"], err_msg=f"{name} {key} {side} {k}")
        g.free()
    ed.free()


@pytest.mark.parametrize("inverse,symmetric", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_inverse_symmetric_file_and_memory(gpu_ctx, golden, tmp_path, inverse, symmetric):
    z = golden("fixture_v1024_e4096")
    p = str(tmp_path / "graph.bin")
    O.write_bin(p, z["edges_src"], z["edges_dst"])
    host = _host_read(p, inverse, symmetric)
    ed = gpu_ctx.edges_read_file(p, inverse=bool(inverse), symmetric=bool(symmetric))
    _same_edges(ed, host)
    ed.free()
    with open(p, "rb") as f:
        raw = f.read()
    em = gpu_ctx.edges_from_records(raw, inverse=bool(inverse), symmetric=bool(symmetric))
    _same_edges(em, host)
    em.free()


def test_weighted_records_special_weights_and_partial_record(gpu_ctx, tmp_path):
    rng = np.random.default_rng(5)
    E = 20000
    s = rng.integers(0, 3000, E).astype(np.uint32)
    d = rng.integers(0, 3000, E).astype(np.uint32)
    w = rng.standard_normal(E).astype(np.float32)
    w[::97] = np.nan
    w[5::89] = -0.0
    w[7::83] = np.float32(1e-42)  # denormal
    w[11::79] = -np.inf
    p = str(tmp_path / "graph.wbin")
    O.write_bin(p, s, d, w)
    with open(p, "ab") as f:
        f.write(b"\x01\x02\x03\x04\x05")  # a partial record: ignored
    for inverse, symmetric in ((0, 0), (1, 1)):
        host = _host_read(p, inverse, symmetric, 1)
        assert host[0].size == E * (2 if symmetric else 1)
        ed = gpu_ctx.edges_read_file(p, weighted_records=True, inverse=bool(inverse), symmetric=bool(symmetric))
        _same_edges(ed, host)
        ed.free()
    # all weights at most zero: the maximum stays 0.0f
    O.write_bin(p, s, d, -np.abs(w))
    host = _host_read(p, 0, 0, 1)
    ed = gpu_ctx.edges_read_file(p, weighted_records=True)
    _same_edges(ed, host)
    assert np.float32(ed.max_weight).view(np.uint32) == 0
    ed.free()


def test_empty_file(gpu_ctx, tmp_path):
    p = str(tmp_path / "empty.bin")
    open(p, "wb").close()
    ed = gpu_ctx.edges_read_file(p)
    s, d, w, N, _ = _host_read(p)
    assert ed.E == 0 == s.size
    ed.free()


@pytest.mark.parametrize("weighted", [False, True])
def test_rmat_s22_against_host_reader_and_host_build(gpu_ctx, tmp_path, weighted):
    s, d, w = gpu_ctx.rmat_edges(22, 16, 24, weight_kind=1 if weighted else 0, weight_seed=24)
    p = str(tmp_path / ("s22.wbin" if weighted else "s22.bin"))
    O.write_bin(p, s, d, w if weighted else None)
    host = _host_read(p, 0, 0, int(weighted))
    ed = gpu_ctx.edges_read_file(p, weighted_records=weighted)
    print(f"[gpu] ingest s22 weighted={weighted}: E={ed.E} N={ed.N}", flush=True)
    _same_edges(ed, host)
    flags = gbx.GRAPH_WEIGHTED if weighted else 0
    g = ed.graph(sort_mode=gbx.SORT_RADIX_SRC_DST, flags=flags)
    ed.free()
    h = gpu_ctx.graph(host[0], host[1], host[2] if weighted else None, sort_mode=gbx.SORT_RADIX_SRC_DST,
                      flags=flags)
    for inv in (False, True):
        a, b = g.export(inverse=inv), h.export(inverse=inv)
        for k in ("out_degree", "in_degree", "edges_idx", "src", "dst") + (("w",) if weighted else ()):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    g.free()
    h.free()


def test_errors(gpu_ctx, tmp_path):
    with pytest.raises(gbx.GbxError, match="open"):
        gpu_ctx.edges_read_file(str(tmp_path / "missing.bin"))
    p = str(tmp_path / "big.bin")
    O.write_bin(p, np.array([0, 1 << 30], np.uint32), np.array([1, 2], np.uint32))
    with pytest.raises(gbx.GbxError, match="2\\^30"):
        gpu_ctx.edges_read_file(p)
    z = np.zeros(8, np.uint8)
    ed = gpu_ctx.edges_from_records(z)  # one record (0, 0): a self loop of vertex 0
    assert (ed.N, ed.E) == (1, 1)
    g = ed.graph()
    with pytest.raises(gbx.GbxError, match="consumed"):
        ed.graph()
    g.free()
    ed.free()
