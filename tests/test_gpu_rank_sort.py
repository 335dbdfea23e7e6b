"""radixSortEdgesByPageRank (reorder.c:165-221) on the GPU and
runPageRankAlgorithm (graphRun.c:933-975) through the adapter: the stable
ascending sort of the labels by FloatToFixed32SORT(rank) = (u32)(rank * 2^20
+ 0.5f) (fixedPoint.h:38) must equal the reference's own sort (compiled from
its sources) and numpy's stable argsort of the same keys, ties included
(keys of normalised ranks below 2^-21 are all 0).  Ranks of 2^12 and above
(keys of 2^32 and above) and negative ranks convert to u32 outside what C
defines: the compiled reference's keys there depend on its code generation,
so those cases are checked against the x86-64 scalar conversion (64-bit
truncation, low 32 bits) only.  PageRank values are in [0, 1]."""
import ctypes as C

import numpy as np
import pytest

import hostapi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ref_sort(ranks, labels):
    R = O.ref(1, 0)
    R.radixSortEdgesByPageRank.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    R.radixSortEdgesByPageRank.restype = C.POINTER(C.c_uint32)
    r = np.ascontiguousarray(ranks, np.float32).copy()
    lab = np.ascontiguousarray(labels, np.uint32).copy()
    out = R.radixSortEdgesByPageRank(r.ctypes.data, lab.ctypes.data, r.size)
    return np.ctypeslib.as_array(out, (r.size,)).copy()


def _keys(ranks):
    r = np.asarray(ranks, np.float32)
    x = (r * np.float32(1048576.0) + np.where(r >= 0, np.float32(0.5), np.float32(-0.5))).astype(np.float32)
    return x.astype(np.int64).astype(np.uint32)


def _cases(rng):
    """(description, ranks, inside the range C defines the conversion for)"""
    n = 200000
    yield "normalised ranks of 1/N scale", (rng.pareto(1.2, n) / n).astype(np.float32), True
    yield "quantised values, many ties", (np.round(rng.random(n) * 50) / 64).astype(np.float32), True
    yield "up to 4095", (rng.random(n) * 4095).astype(np.float32), True
    yield "tiny", np.full(7, 1e-9, np.float32), True
    yield "beyond 2^12: keys wrap at 2^32", (rng.random(n) * 9000).astype(np.float32), False
    yield "some negative", (rng.random(n) - 0.1).astype(np.float32), False


def test_rank_sort_matches_reference_and_stable_argsort(gpu_ctx):
    rng = np.random.default_rng(20261203)
    for what, ranks, defined in _cases(rng):
        print(f"[gpu] rank sort: {what} (n={ranks.size})", flush=True)
        labels = rng.permutation(ranks.size).astype(np.uint32)
        got = gpu_ctx.rank_sort(ranks, labels)
        np.testing.assert_array_equal(got, labels[np.argsort(_keys(ranks), kind="stable")], err_msg=what)
        if defined:
            np.testing.assert_array_equal(got, _ref_sort(ranks, labels), err_msg=what)
    assert gpu_ctx.rank_sort(np.zeros(0, np.float32)).size == 0


def test_run_pagerank_algorithm_sorts_the_ranking(golden, tmp_path):
    L = hostapi.load(False)
    L.runPageRankAlgorithm.restype = C.POINTER(hostapi.PageRankStats)
    L.runPageRankAlgorithm.argtypes = [C.POINTER(hostapi.Arguments), C.c_void_p]
    z = golden("rmat_s12")
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    a = L.argumentsNew()
    keep = path.encode()
    a.contents.fnameb = keep
    a.contents.sort = 1
    a.contents.iterations = 20
    a.contents.epsilon = 0.0
    g = L.graphCSRPreProcessingStep(a)
    N = g.contents.num_vertices
    st = L.runPageRankAlgorithm(a, g)
    ranks = np.ctypeslib.as_array(st.contents.pageRanks, (N,)).copy()
    order = np.ctypeslib.as_array(st.contents.realRanks, (N,)).copy()
    np.testing.assert_array_equal(order, _ref_sort(ranks, np.arange(N, dtype=np.uint32)))
    assert np.all(np.diff(_keys(ranks)[order].astype(np.int64)) >= 0)
    L.freePageRankStats(st)
    a.contents.datastructure = 1  # segments: not implemented, as in the reference
    assert not L.runPageRankAlgorithm(a, g)
    L.graphCSRFree(g)
    L.argumentsFree(a)
