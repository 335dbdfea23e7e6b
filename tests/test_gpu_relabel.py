"""The relabel levels of graphCSRPreProcessingStep on the GPU (lmode,
lmode_l2, lmode_l3; graphCSR.c:268-290, reorderGraphProcess reorder.c:
1556-1616): the new id of every vertex and both CSRs of the relabelled graph
bit-exact with the reference's own pipeline (compiled from its sources), for
every order, with and without -k dedupe, several levels, and a label file;
through the C ABI and through the adapter's graphCSRPreProcessingStep.  A
relabelled graph's fixed-point PageRank is the original's, permuted."""

import numpy as np
import pytest

import graphblox_amd as gbx
import hostapi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LMODES = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14)


def _ref(path, dflag, levels, fnamel=None, sort=1):
    R = O.ref(1, 0)
    g = R.refd_build_relabel(path.encode(), sort, dflag, 0, 0, 1, *levels, fnamel.encode() if fnamel else None)
    N, E = R.refd_num_vertices(g), R.refd_num_edges(g)
    lab = np.zeros(N, np.uint32)
    R.refd_labels(g, lab.ctypes.data, None)
    out = {}
    for side in (0, 1):
        a = [np.zeros(N, np.uint32) for _ in range(3)] + [np.zeros(E, np.uint32) for _ in range(2)]
        R.refd_export(g, side, *[x.ctypes.data for x in a], None)
        out[side] = a
    R.refd_free(g)
    return lab, out


def _check(g, lab, arrays, what):
    np.testing.assert_array_equal(g.labels, lab, err_msg=what)
    for side in (0, 1):
        e = g.export(inverse=bool(side))
        od, ind, idx, src, dst = arrays[side]
        np.testing.assert_array_equal(e["out_degree"], od, err_msg=what)
        np.testing.assert_array_equal(e["edges_idx"], idx, err_msg=what)
        np.testing.assert_array_equal(e["src"], src, err_msg=what)
        np.testing.assert_array_equal(e["dst"], dst, err_msg=what)


@pytest.mark.parametrize("name", ["rmat_s10", "fixture_v1024_e4096", "rmat_s12"])
@pytest.mark.parametrize("dflag", [0, 1])
def test_relabel_matches_reference(gpu_ctx, golden, tmp_path, name, dflag):
    z = golden(name)
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    flags = gbx.GRAPH_DEDUPE if dflag else 0
    for lm in LMODES:
        print(f"[gpu] relabel {name} dflag {dflag} lmode {lm}", flush=True)
        lab, arrays = _ref(path, dflag, (lm, 0, 0))
        g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, flags=flags, lmode=(lm, 0, 0))
        _check(g, lab, arrays, f"lmode {lm}")


def test_relabel_levels_and_file(gpu_ctx, golden, tmp_path):
    z = golden("rmat_s12")
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    for levels in ((4, 1, 8), (6, 0, 0), (0, 3, 0), (13, 9, 7)):
        for sort in (0, 1):
            lab, arrays = _ref(path, 0, levels, sort=sort)
            g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=sort, lmode=levels)
            _check(g, lab, arrays, f"levels {levels} sort {sort}")
    N = int(z["m1_N"])
    perm = np.random.default_rng(5).permutation(N).astype(np.uint32)
    fl = str(tmp_path / "labels.txt")
    np.savetxt(fl, perm, fmt="%u")
    lab, arrays = _ref(path, 0, (11, 0, 0), fnamel=fl)
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, lmode=(11, 0, 0), file_labels=perm)
    _check(g, lab, arrays, "file labels")
    with pytest.raises(gbx.GbxError, match="label"):
        gpu_ctx.graph(z["edges_src"], z["edges_dst"], lmode=(11, 0, 0), file_labels=perm[:10])
    bad = perm.copy()
    bad[3] = N
    with pytest.raises(gbx.GbxError, match="label"):
        gpu_ctx.graph(z["edges_src"], z["edges_dst"], lmode=(11, 0, 0), file_labels=bad)


def test_relabelled_pagerank_is_permuted(gpu_ctx, golden):
    """Renaming vertices does not change the fixed-point PageRank (exact
    integer sums): r_relabelled[label[v]] == r[v]."""
    z = golden("rmat_s12")
    g0 = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    r0, _ = g0.pagerank(iterations=20, epsilon=0.0, mode=gbx.PR_FIXED32)
    for lm in (1, 4, 6):
        g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, lmode=(lm, 0, 0))
        r, _ = g.pagerank(iterations=20, epsilon=0.0, mode=gbx.PR_FIXED32)
        np.testing.assert_array_equal(r[g.labels], r0)


def test_relabel_through_reference_api(golden, tmp_path):
    L = hostapi.load(False)
    z = golden("rmat_s10")
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    for levels, dflag in (((4, 0, 0), 0), ((1, 8, 0), 1)):
        a = L.argumentsNew()
        keep = path.encode()
        a.contents.fnameb = keep
        a.contents.sort = 1
        a.contents.dflag = dflag
        a.contents.lmode, a.contents.lmode_l2, a.contents.lmode_l3 = levels
        g = L.graphCSRPreProcessingStep(a)
        gc = g.contents
        N, E = gc.num_vertices, gc.num_edges
        lab, arrays = _ref(path, dflag, levels)
        el = gc.sorted_edges_array.contents
        np.testing.assert_array_equal(np.ctypeslib.as_array(el.label_array, (N,)), lab)
        inv = np.ctypeslib.as_array(el.inverse_label_array, (N,))
        np.testing.assert_array_equal(inv[lab], np.arange(N))
        for side, (v, e) in enumerate(((gc.vertices, gc.sorted_edges_array),
                                       (gc.inverse_vertices, gc.inverse_sorted_edges_array))):
            od, ind, idx, src, dst = arrays[side]
            np.testing.assert_array_equal(np.ctypeslib.as_array(v.contents.out_degree, (N,)), od)
            np.testing.assert_array_equal(np.ctypeslib.as_array(v.contents.edges_idx, (N,)), idx)
            np.testing.assert_array_equal(np.ctypeslib.as_array(e.contents.edges_array_src, (E,)), src)
            np.testing.assert_array_equal(np.ctypeslib.as_array(e.contents.edges_array_dest, (E,)), dst)
        L.graphCSRFree(g)
        L.argumentsFree(a)
