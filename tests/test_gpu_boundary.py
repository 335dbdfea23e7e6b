"""The drop-in boundary with GraphCSRs this library did not build, the
DIRECTED=0 layout, the upload checks and the hipGraph cache:

* a GraphCSR assembled from the reference's own CSR arrays (as
  graphCSRPreProcessingStep of the reference builds them, graphCSR.c:
  243-327) runs pageRankPullGraphCSR to the reference's ranks;
* a different graph at the same GraphCSR address, and arrays edited in
  place, are uploaded again (fingerprint, graphbloxHipForget) instead of
  running on the stale device copy;
* the DIRECTED=0 adapter (GraphCSR without inverse_*, graphCSR.h:26-30) pulls
  over the forward rows (pageRank.c:269-272, SPMV.c:141-147): its ranks,
  SpMV and fixed-point variants against the reference's DIRECTED=0 build;
* gbx_graph_from_csr rejects vertex arrays that disagree with the edges;
* gbx_pr_launch_graph replays keep the teleport given to the last reset."""
import ctypes as C

import numpy as np
import pytest

import graphblox_amd as gbx
import hostapi
from conftest import l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _csr(z, tag, side):
    return {k: z[f"{tag}_{side}_{k}"].copy() for k in ("out_degree", "in_degree", "edges_idx", "src", "dst")}


def _pr(L, ptr, iterations=20):
    a = L.argumentsNew()
    a.contents.iterations = iterations
    a.contents.epsilon = 0.0
    st = L.pageRankPullGraphCSR(a, ptr)
    assert st
    r = np.ctypeslib.as_array(st.contents.pageRanks, (st.contents.num_vertices,)).copy()
    L.freePageRankStats(st)
    L.argumentsFree(a)
    return r


def test_adapter_runs_on_a_graph_built_by_the_reference(golden):
    """The reference's own forward and transpose arrays in a GraphCSR of the
    reference's layout, handed to pageRankPullGraphCSR unchanged."""
    L = hostapi.load(False)
    z = golden("rmat_s12")
    N, E = int(z["m1_N"]), int(z["m1_E"])
    hg = hostapi.HostGraph(L, N, E, _csr(z, "m1", "fwd"), _csr(z, "m1", "inv"))
    assert l1(_pr(L, hg.ptr), z["pr_r20"]) <= TOL
    L.graphbloxHipShutdown()


def _relabelled(z, tag, perm):
    """The graph with vertex v renamed perm[v]: CSRs rebuilt by the oracle."""
    N = int(z[f"{tag}_N"])
    s, d = z[f"{tag}_fwd_src"], z[f"{tag}_fwd_dst"]
    fwd, inv = O.build_graph(perm[s].astype(np.uint32), perm[d].astype(np.uint32), None, N, mode=1)
    pick = lambda c: dict(out_degree=c.out_degree, in_degree=c.in_degree, edges_idx=c.edges_idx, src=c.src, dst=c.dst)
    return pick(fwd), pick(inv)


def test_stale_device_copy_is_not_used(golden):
    """A new graph built at the address of a freed one, and arrays edited in
    place, give the new graph's ranks."""
    L = hostapi.load(False)
    z10, z12 = golden("rmat_s10"), golden("rmat_s12")
    hg = hostapi.HostGraph(L, int(z10["m1_N"]), int(z10["m1_E"]), _csr(z10, "m1", "fwd"), _csr(z10, "m1", "inv"))
    assert l1(_pr(L, hg.ptr), z10["pr_r20"]) <= TOL
    addr = hg.ptr
    # the same struct now describes another graph (as after a free and a new build at that address)
    hg.set(int(z12["m1_N"]), int(z12["m1_E"]), _csr(z12, "m1", "fwd"), _csr(z12, "m1", "inv"))
    assert hg.ptr == addr
    assert l1(_pr(L, hg.ptr), z12["pr_r20"]) <= TOL
    # the arrays edited in place: the same graph with its vertices renamed
    N = int(z12["m1_N"])
    perm = np.random.default_rng(3).permutation(N).astype(np.uint32)
    fwd, inv = _relabelled(z12, "m1", perm)
    for side, new in ((hg.g.vertices.contents, fwd), (hg.g.inverse_vertices.contents, inv)):
        for nm in ("out_degree", "in_degree", "edges_idx"):
            np.ctypeslib.as_array(getattr(side, nm), (N,))[:] = new[nm]
    E = int(z12["m1_E"])
    for side, new in ((hg.g.sorted_edges_array.contents, fwd), (hg.g.inverse_sorted_edges_array.contents, inv)):
        np.ctypeslib.as_array(side.edges_array_src, (E,))[:] = new["src"]
        np.ctypeslib.as_array(side.edges_array_dest, (E,))[:] = new["dst"]
    want = np.empty(N, np.float32)
    want[perm] = z12["pr_r20"]
    assert l1(_pr(L, hg.ptr), want) <= TOL  # noticed through the sampled entries
    # an edit the samples cannot see (entry 1 of the transpose's column array
    # lies between the first two samples): graphbloxHipForget uploads it
    fo, io = _csr(z12, "m1", "fwd"), _csr(z12, "m1", "inv")
    hg.set(N, E, fo, io)
    r_before = _pr(L, hg.ptr)
    cols = np.ctypeslib.as_array(hg.g.inverse_sorted_edges_array.contents.edges_array_dest, (E,))
    od = fo["out_degree"].astype(np.float64)
    c = np.where(od > 0, z12["pr_r20"] / np.maximum(od, 1), 0.0)
    order = np.argsort(c)
    cols[1] = np.uint32(order[-1] if order[-1] != cols[1] else order[-2])  # the largest other contribution
    np.testing.assert_array_equal(_pr(L, hg.ptr), r_before)  # the fingerprint missed this edit
    L.graphbloxHipForget(hg.ptr)
    r_after = _pr(L, hg.ptr)
    io["dst"] = cols.copy()
    fwd_o = O.OracleCSR(N, E, fo["out_degree"], fo["in_degree"], fo["edges_idx"], fo["src"], fo["dst"], None)
    inv_o = O.OracleCSR(N, E, io["out_degree"], io["in_degree"], io["edges_idx"], io["src"], io["dst"], None)
    want = O.pr_pull_f32(fwd_o, inv_o, 20, 0.0, nthreads=4)["ranks"]
    assert l1(r_after, want) <= TOL and not np.array_equal(r_after, r_before)
    L.graphbloxHipShutdown()


def test_directed0_adapter_matches_the_reference_directed0_build(golden, tmp_path):
    """graphCSRPreProcessingStep and the pulls of the DIRECTED=0 adapter
    against the reference built with DIRECTED 0 (its shipped setting)."""
    L = hostapi.load(False, directed=False)
    L.SPMVGraphCSR.restype = C.POINTER(hostapi.SPMVStats)
    z = golden("rmat_s12")
    ref = golden("reference_api")
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    a = L.argumentsNew()
    keep = path.encode()
    a.contents.fnameb = keep
    a.contents.sort = 1
    a.contents.iterations = 20
    a.contents.epsilon = 0.0
    g = L.graphCSRPreProcessingStep(a)
    gc = g.contents
    N, E = gc.num_vertices, gc.num_edges
    assert (N, E) == (int(z["m1_N"]), int(z["m1_E"]))
    # the forward CSR as the reference builds it; in_degree stays zero without DIRECTED (vertex.c:291-311)
    v, e = gc.vertices.contents, gc.sorted_edges_array.contents
    np.testing.assert_array_equal(np.ctypeslib.as_array(v.out_degree, (N,)), z["m1_fwd_out_degree"])
    np.testing.assert_array_equal(np.ctypeslib.as_array(v.edges_idx, (N,)), z["m1_fwd_edges_idx"])
    assert not np.ctypeslib.as_array(v.in_degree, (N,)).any()
    np.testing.assert_array_equal(np.ctypeslib.as_array(e.edges_array_dest, (E,)), z["m1_fwd_dst"])
    for code in (0, 1, 2, 3, 4, 5, 9):
        a.contents.pushpull = code
        st = L.pageRankGraphCSR(a, g)
        assert st, code
        r = np.ctypeslib.as_array(st.contents.pageRanks, (N,)).copy()
        want = ref[f"d0/pr/code{code}/r20"]
        if code in (0, 1):  # float sums: the pull iteration's tolerance
            assert l1(r, want) <= TOL
        elif code == 5:  # float-atomic sums of the 32-bit quantized terms
            assert l1(r, want) <= 1e-6
        else:
            np.testing.assert_array_equal(r, want, err_msg=f"code {code}")
        L.freePageRankStats(st)
    a.contents.pushpull = 6
    assert not L.pageRankGraphCSR(a, g)  # the data-driven variants: DIRECTED=1 builds only
    a.contents.iterations = 3
    for code in (0, 1, 2, 3):
        a.contents.pushpull = code
        st = L.SPMVGraphCSR(a, g) if code else L.runSPMVAlgorithm(a, g)
        assert st, code
        y = np.ctypeslib.as_array(st.contents.vector_output, (N,)).copy()
        want = ref[f"d0/spmv/w0/code{code}"]
        if code >= 2:
            np.testing.assert_array_equal(y, want, err_msg=f"code {code}")
        else:
            assert l1(y, want) <= 1e-6 * float(np.abs(want.astype(np.float64)).sum())
        L.freeSPMVStats(st)
    L.graphCSRFree(g)
    # a forward-only GraphCSR of the reference's arrays (the transpose is built on the device)
    hg = hostapi.HostGraph(L, N, E, _csr(z, "m1", "fwd"))
    assert l1(_pr(L, hg.ptr), z["pr_fwd_r20"]) <= TOL
    L.graphbloxHipShutdown()
    L.argumentsFree(a)


def test_directed0_weighted_spmv(golden, tmp_path):
    L = hostapi.load(True, directed=False)
    z = golden("rmat_s12")
    ref = golden("reference_api")
    path = str(tmp_path / "g.wbin")
    O.write_bin(path, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
    a = L.argumentsNew()
    keep = path.encode()
    a.contents.fnameb = keep
    a.contents.sort = 1
    a.contents.weighted = 1
    a.contents.iterations = 3
    g = L.graphCSRPreProcessingStep(a)
    N = g.contents.num_vertices
    for code in (0, 1, 2, 3):
        a.contents.pushpull = code
        st = L.SPMVGraphCSR(a, g)
        assert st, code
        y = np.ctypeslib.as_array(st.contents.vector_output, (N,)).copy()
        want = ref[f"d0/spmv/w1/code{code}"]
        if code >= 2:
            np.testing.assert_array_equal(y, want, err_msg=f"code {code}")
        else:
            assert l1(y, want) <= 1e-6 * float(np.abs(want.astype(np.float64)).sum())
        L.freeSPMVStats(st)
    L.graphCSRFree(g)
    L.graphbloxHipShutdown()
    L.argumentsFree(a)


def test_upload_rejects_inconsistent_vertex_arrays(gpu_ctx, golden):
    """Vertex arrays that disagree with the sorted edges are refused before
    any kernel indexes with them."""
    z = golden("rmat_s10")
    fwd = {k: z[f"m1_fwd_{k}"].copy() for k in ("out_degree", "edges_idx", "src", "dst")}
    inv = {k: z[f"m1_inv_{k}"].copy() for k in ("out_degree", "edges_idx", "src", "dst")}
    g = gpu_ctx.graph_from_csr(fwd, inv)  # consistent: accepted
    g.free()
    v = int(np.flatnonzero(inv["out_degree"] > 1)[0])
    cases = []
    bad = {k: a.copy() for k, a in inv.items()}
    bad["edges_idx"][v] += 1  # v's first edge falls before its range
    cases.append(("edges_idx past the first edge", fwd, bad))
    bad = {k: a.copy() for k, a in inv.items()}
    bad["out_degree"][v] += 1  # degrees no longer sum to E
    cases.append(("a degree too large", fwd, bad))
    bad = {k: a.copy() for k, a in fwd.items()}
    u = int(np.flatnonzero(fwd["out_degree"] > 0)[-1])
    bad["out_degree"][u] -= 1  # u's last edge falls outside its range
    cases.append(("a degree too small", bad, inv))
    bad = {k: a.copy() for k, a in fwd.items()}
    bad["edges_idx"][u] = np.uint32(0xFFFFFFF0)
    cases.append(("edges_idx beyond E", bad, inv))
    for what, f, i in cases:
        with pytest.raises(gbx.GbxError, match="out_degree / edges_idx"):
            gpu_ctx.graph_from_csr(f, i)
        print(f"[gpu] rejected: {what}", flush=True)


@pytest.mark.parametrize("method", [gbx.METHOD_PULL, gbx.METHOD_BINNED])
def test_launch_graph_follows_the_teleport_of_each_reset(gpu_ctx, method):
    """reset(NULL) then a captured replay, reset(t) then a replay, and back:
    each bit-identical with eager launches after the same reset."""
    g = gpu_ctx.rmat_graph(14, 16, 11)
    N = g.N
    rng = np.random.default_rng(4)
    t = rng.random(N).astype(np.float64)
    t = (t / t.sum()).astype(np.float32)
    kw = dict(iterations=12, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE, method=method)
    eager = g.session(**kw)
    replay = g.session(**kw)
    for tele in (None, t, None, t):
        eager.reset(tele)
        eager.launch(6)
        eager.launch(6)
        want, _ = eager.read()
        replay.reset(tele)
        replay.launch_graph(6)
        replay.launch_graph(6)
        got, _ = replay.read()
        np.testing.assert_array_equal(got, want)
    eager.free()
    replay.free()
    g.free()
