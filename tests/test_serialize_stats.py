"""The reference's CSR files (writeToBinFileGraphCSR, readFromBinFileGraphCSR,
writetoTextFilesGraphCSR; graphCSR.c:457-744) and PageRank top-k similarity
statistics (collectStatsPageRank_topK / collectStatsPageRank, graphStats.c:
589-1010) through the adapter, against the reference's own code compiled from
its sources (oracle/_ref): byte-identical files, the same structures read
back by either reader, identical statistics.  All host code (no GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

import hostapi
from oracle import oracle as O


class CorrStats(C.Structure):
    _fields_ = [("levenshtein_distance", C.c_uint32), ("float_Kendall", C.c_float), ("real_Kendall", C.c_float),
                ("intersection", C.c_uint32), ("mismatch", C.c_uint32), ("avg_error_float", C.c_double),
                ("avg_error_relative", C.c_double)]


def _bind_io(L, graph_ptr):
    L.writeToBinFileGraphCSR.argtypes = [C.c_char_p, C.c_void_p]
    L.writetoTextFilesGraphCSR.argtypes = [C.c_char_p, C.c_void_p]
    L.readFromBinFileGraphCSR.argtypes = [C.c_char_p]
    L.readFromBinFileGraphCSR.restype = graph_ptr
    return L


def _bind_stats(L):
    pS = C.POINTER(hostapi.PageRankStats)
    u32p = C.POINTER(C.c_uint32)
    L.collectStatsPageRank_topK.argtypes = [pS, pS, u32p, u32p, u32p, C.c_uint32, C.c_uint32, C.c_void_p,
                                            C.c_uint32]
    L.collectStatsPageRank_topK.restype = CorrStats
    L.collectStatsPageRank.argtypes = [C.POINTER(hostapi.Arguments), pS, pS, C.c_uint32]
    return L


def _edges_file(tmp_path, z, weighted):
    p = str(tmp_path / ("g.wbin" if weighted else "g.bin"))
    if weighted:
        O.write_bin(p, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
    else:
        O.write_bin(p, z["edges_src"], z["edges_dst"])
    return p


def _export(R, g, N, E, sides=(0, 1)):
    out = {}
    for side in sides:
        a = [np.zeros(N, np.uint32) for _ in range(3)] + [np.zeros(E, np.uint32) for _ in range(2)] + \
            [np.zeros(E, np.float32)]
        R.refd_export(g, side, *[x.ctypes.data for x in a])
        out[side] = a
    return out


def _arr(ptr, n, dt):
    return np.ctypeslib.as_array(ptr, (n,)).astype(dt, copy=True)


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("name", ["rmat_s10", "rmat_s12"])
def test_csr_files_match_the_reference(golden, tmp_path, name, weighted, directed):
    """In each (DIRECTED, WEIGHTED) layout: the reference's own build of that
    layout writes the GraphCSR, the adapter of the same layout writes it too,
    and each reads the other's file."""
    z = golden(name)
    R = _bind_io(O.ref(int(directed), int(weighted)), C.c_void_p)
    L = hostapi.load(weighted, directed)
    _bind_io(L, C.POINTER(L.GraphCSR))
    sides = (0, 1) if directed else (0,)
    g = R.refd_build(_edges_file(tmp_path, z, weighted).encode(), 1, 0, 0, int(weighted), 1)
    N, E = R.refd_num_vertices(g), R.refd_num_edges(g)
    ref_base, mine_base = str(tmp_path / "ref"), str(tmp_path / "mine")
    R.writeToBinFileGraphCSR(ref_base.encode(), g)
    L.writeToBinFileGraphCSR(mine_base.encode(), g)
    ref_bytes = open(ref_base + ".csr", "rb").read()
    assert open(mine_base + ".csr", "rb").read() == ref_bytes
    assert len(ref_bytes) == 4 * (2 + int(weighted) + len(sides) * (3 * N + (3 if weighted else 2) * E))
    R.writetoTextFilesGraphCSR(ref_base.encode(), g)
    L.writetoTextFilesGraphCSR(mine_base.encode(), g)
    exts = [".out_degree", ".in_degree", ".edges_idx", ".edges_array_src", ".edges_array_dest"]
    for ext in exts + ([".edges_array_weight"] if weighted else []):
        assert open(mine_base + ext).read() == open(ref_base + ext).read(), ext
    # read back by the adapter: the same arrays, and the same file when written again
    ref_arrays = _export(R, g, N, E, sides)
    gm = L.readFromBinFileGraphCSR((mine_base + ".csr").encode())
    c = gm.contents
    assert (c.num_vertices, c.num_edges) == (N, E)
    pairs = [(c.vertices, c.sorted_edges_array)]
    if directed:
        pairs.append((c.inverse_vertices, c.inverse_sorted_edges_array))
    for side, (v, e) in enumerate(pairs):
        od, ind, idx, src, dst, w = ref_arrays[side]
        np.testing.assert_array_equal(_arr(v.contents.out_degree, N, np.uint32), od)
        np.testing.assert_array_equal(_arr(v.contents.in_degree, N, np.uint32), ind)
        np.testing.assert_array_equal(_arr(v.contents.edges_idx, N, np.uint32), idx)
        np.testing.assert_array_equal(_arr(e.contents.edges_array_src, E, np.uint32), src)
        np.testing.assert_array_equal(_arr(e.contents.edges_array_dest, E, np.uint32), dst)
        if weighted:
            np.testing.assert_array_equal(_arr(e.contents.edges_array_weight, E, np.float32), w)
    again = str(tmp_path / "again")
    L.writeToBinFileGraphCSR(again.encode(), gm)
    assert open(again + ".csr", "rb").read() == ref_bytes
    L.graphCSRFree(gm)
    # and the reference's reader takes the adapter's file
    gr = R.readFromBinFileGraphCSR((mine_base + ".csr").encode())
    back = _export(R, gr, N, E, sides)
    for side in sides:
        for a, b in zip(back[side], ref_arrays[side]):
            np.testing.assert_array_equal(a, b)
    R.refd_free(gr)
    R.refd_free(g)


def test_csr_reader_exits_on_a_missing_file(tmp_path):
    import subprocess
    import sys

    from conftest import ROOT
    code = ("import sys; sys.path[:0] = [%r, %r]; import ctypes as C, hostapi; L = hostapi.load(False); "
            "L.readFromBinFileGraphCSR.argtypes = [C.c_char_p]; L.readFromBinFileGraphCSR(b'%s')"
            % (os.path.join(ROOT, "tests"), ROOT, str(tmp_path / "none.csr")))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode == 1 and "open" in r.stderr


def _stats(ranks, realRanks, keep):
    """A PageRankStats over numpy arrays (kept alive in `keep`)."""
    ranks = np.ascontiguousarray(ranks, np.float32)
    realRanks = np.ascontiguousarray(realRanks, np.uint32)
    keep += [ranks, realRanks]
    s = hostapi.PageRankStats()
    s.damp, s.base_pr, s.iterations, s.num_vertices = 0.85, 0.15, 7, ranks.size
    s.realRanks = realRanks.ctypes.data_as(C.POINTER(C.c_uint32))
    s.pageRanks = ranks.ctypes.data_as(C.POINTER(C.c_float))
    s.time_total, s.error_total = 0.25, 0.0
    return s


def _two_runs(n, seed):
    """A reference ranking and a perturbed one, with tied values."""
    rng = np.random.default_rng(seed)
    ref = np.round(rng.pareto(1.5, n) * 50) / 1000 + 1e-4
    run = ref * (1 + rng.normal(0, 0.02, n))
    run[rng.choice(n, n // 20, replace=False)] = 0.0
    ref = ref.astype(np.float32)
    run = run.astype(np.float32)
    return ref, run, np.argsort(ref, kind="stable"), np.argsort(run, kind="stable")


def test_topk_statistics_match_the_reference():
    L, R = _bind_stats(hostapi.load(False)), _bind_stats(O.ref(1, 0))
    u32p = C.POINTER(C.c_uint32)
    for seed, n in ((1, 3000), (2, 1200), (3, 40)):
        keep = []
        ref, run, ref_order, run_order = _two_runs(n, seed)
        a, b = _stats(ref, ref_order, keep), _stats(run, run_order, keep)
        run_pos = np.empty(n, np.uint32)
        run_pos[run_order] = np.arange(n, dtype=np.uint32)
        ref_pos = np.empty(n, np.uint32)
        ref_pos[ref_order] = np.arange(n, dtype=np.uint32)
        total = ref_order.astype(np.uint32)
        args = [x.ctypes.data_as(u32p) for x in (total, ref_pos, run_pos)]
        for topk in (1, 30, 100, 300, 1000, 2500, 5000):
            for nv in (n, max(1, n - 700)):
                got = L.collectStatsPageRank_topK(C.byref(a), C.byref(b), *args, topk, nv, None, 0)
                want = R.collectStatsPageRank_topK(C.byref(a), C.byref(b), *args, topk, nv, None, 0)
                for f, _ in CorrStats._fields_:
                    assert getattr(got, f) == getattr(want, f), (seed, topk, nv, f)


def test_topk_edit_distance_follows_the_reference_rule():
    """The reference's minimum of three picks the third cost on a tie of the
    first two, so its distance can exceed the textbook edit distance; these
    sequences show it, and the adapter reports the reference's value."""
    L, R = _bind_stats(hostapi.load(False)), _bind_stats(O.ref(1, 0))
    u32p = C.POINTER(C.c_uint32)

    def textbook(x, y):
        d = np.arange(len(x) + 1)
        for j in range(1, len(y) + 1):
            prev, d = d, np.zeros_like(d)
            d[0] = j
            for i in range(1, len(x) + 1):
                d[i] = min(prev[i] + 1, d[i - 1] + 1, prev[i - 1] + (x[i - 1] != y[j - 1]))
        return int(d[-1])

    differs = 0
    rng = np.random.default_rng(100)
    for _ in range(600):
        n = 12
        ref_order = rng.permutation(n).astype(np.uint32)
        run_order = rng.permutation(n).astype(np.uint32)
        keep = []
        ranks = np.linspace(1, 2, n).astype(np.float32)
        a, b = _stats(ranks, ref_order, keep), _stats(ranks, run_order, keep)
        pos = np.zeros(n, np.uint32)
        args = [x.ctypes.data_as(u32p) for x in (ref_order, pos, pos)]
        got = L.collectStatsPageRank_topK(C.byref(a), C.byref(b), *args, n, n, None, 0)
        want = R.collectStatsPageRank_topK(C.byref(a), C.byref(b), *args, n, n, None, 0)
        assert got.levenshtein_distance == want.levenshtein_distance
        differs += got.levenshtein_distance != textbook(list(ref_order), list(run_order))
    assert differs > 0


@pytest.mark.parametrize("verbosity", [0, 1, 2])
def test_collect_stats_file_matches_the_reference(tmp_path, verbosity):
    L, R = _bind_stats(hostapi.load(False)), _bind_stats(O.ref(1, 0))
    keep = []
    ref, run, ref_order, run_order = _two_runs(2500, 7)
    a, b = _stats(ref, ref_order, keep), _stats(run, run_order, keep)
    outs = []
    for lib, tag in ((R, "ref"), (L, "mine")):
        A = hostapi.Arguments()
        base = str(tmp_path / tag).encode()
        keep.append(base)
        A.fnameb = base
        A.algorithm, A.datastructure, A.pushpull = 1, 0, 4
        A.binSize, A.verbosity, A.pre_numThreads = 1000, verbosity, 8
        lib.collectStatsPageRank(C.byref(A), C.byref(a), C.byref(b), 3)
        outs.append(open(str(tmp_path / f"{tag}_1_0_3_4.stats")).read())
    assert outs[0] == outs[1]
