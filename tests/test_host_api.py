"""The reference host API served by the HIP backend (include/graphblox_hip.h),
driven the way the reference's own host code drives it (graphRun.c,
tests_c/test_match.c): argumentsNew -> graphCSRPreProcessingStep ->
pageRankGraphCSR / SPMVGraphCSR -> free."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, l1
from oracle import oracle as O


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:struct \w+ \*|void|int)\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("weighted", [False, True])
def test_adapter_exports_every_declared_function(weighted):
    import hostapi

    L = hostapi.load(weighted)
    names = _declared("graphblox_hip.h")
    assert {"pageRankPullGraphCSR", "SPMVPullGraphCSR", "graphCSRPreProcessingStep"} <= set(names)
    assert [n for n in names if not hasattr(L, n)] == []


def test_reader_n_rule_and_arrays(golden, tmp_path):
    """readEdgeListsbin (CPU only): N = max id + 1, symmetric and inverse."""
    import hostapi

    L = hostapi.load(False)
    z = golden("fixture_v500_e500")
    p = str(tmp_path / "g.bin")
    O.write_bin(p, z["edges_src"], z["edges_dst"])
    s, d, _, N = O.read_bin(p)
    for inverse, symmetric in ((0, 0), (1, 0), (0, 1)):
        e = L.readEdgeListsbin(p.encode(), inverse, symmetric, 0).contents
        E = s.size * (2 if symmetric else 1)
        assert (e.num_edges, e.num_vertices) == (E, N)
        gs = np.ctypeslib.as_array(e.edges_array_src, (E,))
        gd = np.ctypeslib.as_array(e.edges_array_dest, (E,))
        es, ed = (d, s) if inverse else (s, d)
        if symmetric:
            es, ed = np.concatenate([s, d]), np.concatenate([d, s])
        np.testing.assert_array_equal(gs, es)
        np.testing.assert_array_equal(gd, ed)
        L.freeEdgeList(C.pointer(e))


def test_reader_missing_file_exits_like_reference(tmp_path):
    code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r); import hostapi; "
            "hostapi.load(False).readEdgeListsbin(b'/nonexistent/graph.bin', 0, 0, 0)"
            % (os.path.join(ROOT, "tests"), ROOT))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode == 1
    assert "open" in r.stderr


def _write(tmp_path, z, weighted=False):
    p = str(tmp_path / ("g.wbin" if weighted else "g.bin"))
    if weighted:
        O.write_bin(p, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
    else:
        O.write_bin(p, z["edges_src"], z["edges_dst"])
    return p


def _arr(ptr, n, dt=np.uint32):
    return np.ctypeslib.as_array(ptr, (n,)).astype(dt, copy=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["fixture_v1024_e4096", "rmat_s10", "rmat_s12"])
@pytest.mark.parametrize("sort,tag,dflag", [(1, "m1", 0), (2, "m2", 0), (3, "m1", 0), (1, "k", 1)])
def test_preprocessing_and_pagerank_through_reference_api(golden, tmp_path, name, sort, tag, dflag):
    import hostapi

    L = hostapi.load(False)
    z = golden(name)
    path = _write(tmp_path, z)
    a = L.argumentsNew()
    keep = path.encode()
    a.contents.fnameb = keep
    a.contents.sort = sort
    a.contents.dflag = dflag
    a.contents.iterations = 20
    a.contents.epsilon = 0.0
    g = L.graphCSRPreProcessingStep(a)
    gc = g.contents
    assert (gc.num_vertices, gc.num_edges) == (int(z[f"{tag}_N"]), int(z[f"{tag}_E"]))
    N, E = gc.num_vertices, gc.num_edges
    for side, v, e in (("fwd", gc.vertices, gc.sorted_edges_array), ("inv", gc.inverse_vertices,
                                                                      gc.inverse_sorted_edges_array)):
        np.testing.assert_array_equal(_arr(v.contents.out_degree, N), z[f"{tag}_{side}_out_degree"])
        np.testing.assert_array_equal(_arr(v.contents.in_degree, N), z[f"{tag}_{side}_in_degree"])
        np.testing.assert_array_equal(_arr(v.contents.edges_idx, N), z[f"{tag}_{side}_edges_idx"])
        np.testing.assert_array_equal(_arr(e.contents.edges_array_src, E), z[f"{tag}_{side}_src"])
        np.testing.assert_array_equal(_arr(e.contents.edges_array_dest, E), z[f"{tag}_{side}_dst"])
    if tag == "m1":
        st = L.pageRankGraphCSR(a, g)
        assert st
        s = st.contents
        assert s.iterations == 20 and s.num_vertices == N
        assert s.damp == np.float32(0.85) and s.base_pr == np.float32(1.0) - np.float32(0.85)
        assert l1(_arr(s.pageRanks, N, np.float32), z["pr_r20"]) <= 1e-5
        np.testing.assert_array_equal(_arr(s.realRanks, N), np.arange(N))
        L.freePageRankStats(st)
        a.contents.iterations = 100
        a.contents.epsilon = 1e-4
        st = L.pageRankPullGraphCSR(a, g)
        assert abs(st.contents.iterations - int(z["pr_eps_iters"])) <= 1
        L.freePageRankStats(st)
        a.contents.iterations = 3
        sp = L.SPMVGraphCSR(a, g)
        np.testing.assert_array_equal(_arr(sp.contents.vector_input, N, np.float32), z["spmv_x"])
        y = _arr(sp.contents.vector_output, N, np.float32)
        assert l1(y, z["spmv_y3"]) <= 1e-6 * float(np.abs(z["spmv_y3"]).sum())
        L.freeSPMVStats(sp)
        # the reference's variants by pushpull code: push sums as pull; fixed point bit-exact
        v = golden(f"variants_{name}")
        a.contents.iterations = 20
        a.contents.epsilon = 0.0
        for code in (1, 2, 3, 4, 5, 6, 9, 10, 11, 12, 13):
            a.contents.pushpull = code
            st = L.pageRankGraphCSR(a, g)
            assert st, code
            r = _arr(st.contents.pageRanks, N, np.float32)
            if code in (1, 6):  # float sums: the pull iteration's tolerance
                assert l1(r, v[f"pp{code}_r20"]) <= 1e-5
            elif code == 5:  # float-atomic sums of the 32-bit quantized terms (7e-8 from code 4 in the reference)
                assert l1(r, v["pp5_r20"]) <= 1e-6
            else:
                np.testing.assert_array_equal(r, v[f"pp{code}_r20"])
            L.freePageRankStats(st)
        # data-driven push and pull-push: their limits as epsilon shrinks
        fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, N, mode=1)
        rp = O.row_ptr_from(inv)
        lim = {7: O.dd_push_limit(N, fwd.out_degree, rp, inv.dst),
               8: O.pr_f64(N, fwd.out_degree, rp, inv.dst, mode=0, iterations=3000, epsilon=1e-15)[0]}
        a.contents.iterations = 3000
        a.contents.epsilon = 1e-7
        for code in (7, 8):
            a.contents.pushpull = code
            st = L.pageRankGraphCSR(a, g)
            assert st, code
            # the pull-push form's stop rule bounds its pushed residuals only (tests/test_gpu_dd_push.py)
            assert l1(_arr(st.contents.pageRanks, N, np.float32), lim[code]) <= (1e-7 / 0.15 + 1e-6 if code == 7
                                                                                  else 1e-4)
            L.freePageRankStats(st)
        a.contents.pushpull = 0
    L.graphCSRFree(g)
    L.argumentsFree(a)


@pytest.mark.gpu
def test_weighted_build_through_reference_api(golden, tmp_path):
    import hostapi

    L = hostapi.load(True)
    z = golden("rmat_s10")
    path = _write(tmp_path, z, weighted=True)
    a = L.argumentsNew()
    keep = path.encode()
    a.contents.fnameb = keep
    a.contents.weighted = 1
    a.contents.sort = 1
    a.contents.iterations = 1
    g = L.graphCSRPreProcessingStep(a)
    gc = g.contents
    N, E = gc.num_vertices, gc.num_edges
    np.testing.assert_array_equal(_arr(gc.inverse_sorted_edges_array.contents.edges_array_weight, E, np.float32),
                                  z["wm1_inv_w"])
    sp = L.SPMVGraphCSR(a, g)
    y = _arr(sp.contents.vector_output, N, np.float32)
    assert l1(y, z["wspmv_y1"]) <= 1e-6 * float(np.abs(z["wspmv_y1"]).sum())
    L.freeSPMVStats(sp)
    L.graphCSRFree(g)
