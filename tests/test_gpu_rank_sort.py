"""radixSortEdgesByPageRank (reorder.c:165-221) on the GPU and
runPageRankAlgorithm (graphRun.c:933-975) through the adapter: the stable
ascending sort of the labels by FloatToFixed32SORT(rank) = (u32)(rank * 2^20
+ 0.5f) (fixedPoint.h:38) must equal the reference's own sort (the digests
tests/golden/make_reference_api_goldens.py recorded from the reference
compiled from its sources) and numpy's stable argsort of the same keys, ties
included (keys of normalised ranks below 2^-21 are all 0).  Ranks of 2^12 and above
(keys of 2^32 and above) and negative ranks convert to u32 outside what C
defines: the compiled reference's keys there depend on its code generation,
so those cases are checked against the x86-64 scalar conversion (64-bit
truncation, low 32 bits) only.  PageRank values are in [0, 1]."""
import ctypes as C

import numpy as np
import pytest

import hostapi
import refapi_cases as RC
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _keys(ranks):
    r = np.asarray(ranks, np.float32)
    x = (r * np.float32(1048576.0) + np.where(r >= 0, np.float32(0.5), np.float32(-0.5))).astype(np.float32)
    return x.astype(np.int64).astype(np.uint32)


def test_rank_sort_matches_reference_and_stable_argsort(gpu_ctx, golden):
    ref = golden("reference_api")
    for i, (what, ranks, labels, defined) in enumerate(RC.rank_sort_cases()):
        print(f"[gpu] rank sort: {what} (n={ranks.size})", flush=True)
        got = gpu_ctx.rank_sort(ranks, labels)
        np.testing.assert_array_equal(got, labels[np.argsort(_keys(ranks), kind="stable")], err_msg=what)
        if defined:
            assert RC.sha(got) == str(ref[f"ranksort/{i}_sha"]), f"{what}: differs from the reference's sort"
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
    # the stable ascending sort of the keys, which the test above pins to the reference's sort
    np.testing.assert_array_equal(order, np.argsort(_keys(ranks), kind="stable").astype(np.uint32))
    assert np.all(np.diff(_keys(ranks)[order].astype(np.int64)) >= 0)
    L.freePageRankStats(st)
    a.contents.datastructure = 1  # segments: not implemented, as in the reference
    assert not L.runPageRankAlgorithm(a, g)
    L.graphCSRFree(g)
    L.argumentsFree(a)
