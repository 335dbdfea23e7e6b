"""The relabel levels of graphCSRPreProcessingStep on the GPU (lmode,
lmode_l2, lmode_l3; graphCSR.c:268-290, reorderGraphProcess reorder.c:
1556-1616): the new id of every vertex and both CSRs of the relabelled graph
bit-exact with the reference's own pipeline, for every order, with and
without -k dedupe, several levels, and a label file; through the C ABI and
through the adapter's graphCSRPreProcessingStep.  The reference's outputs
are the digests (and, for the small graphs, the labels) that
tests/golden/make_reference_api_goldens.py recorded from the reference
compiled from its sources.  A relabelled graph's fixed-point PageRank is the
original's, permuted."""

import numpy as np
import pytest

import graphblox_amd as gbx
import hostapi
import refapi_cases as RC
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_arrays(ref, key, labels, sides, what):
    if f"{key}/labels" in ref:
        np.testing.assert_array_equal(labels, ref[f"{key}/labels"], err_msg=what)
    assert RC.sha(labels) == str(ref[f"{key}/labels_sha"]), f"{what}: labels differ from the reference's"
    for side in (0, 1):
        for nm, a in sides[side].items():
            assert RC.sha(a) == str(ref[f"{key}/{side}/{nm}_sha"]), f"{what}: side {side} {nm} differs"


def _check(ref, key, g, what):
    sides = {}
    for side in (0, 1):
        e = g.export(inverse=bool(side))
        sides[side] = {nm: e[nm] for nm in ("out_degree", "edges_idx", "src", "dst")}
    _check_arrays(ref, key, g.labels, sides, what)


@pytest.mark.parametrize("name", list(RC.RELABEL_GRAPHS))
@pytest.mark.parametrize("dflag", [0, 1])
def test_relabel_matches_reference(gpu_ctx, golden, name, dflag):
    z = golden(name)
    ref = golden("reference_api")
    flags = gbx.GRAPH_DEDUPE if dflag else 0
    for lm in RC.LMODES:
        print(f"[gpu] relabel {name} dflag {dflag} lmode {lm}", flush=True)
        g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, flags=flags, lmode=(lm, 0, 0))
        _check(ref, RC.relabel_key(name, dflag, (lm, 0, 0), 1), g, f"lmode {lm}")
        g.free()


def test_relabel_levels_and_file(gpu_ctx, golden):
    z = golden("rmat_s12")
    ref = golden("reference_api")
    for levels in RC.LEVELS:
        for sort in (0, 1):
            g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=sort, lmode=levels)
            _check(ref, RC.relabel_key("rmat_s12", 0, levels, sort), g, f"levels {levels} sort {sort}")
            g.free()
    N = int(z["m1_N"])
    perm = RC.file_labels(N)
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, lmode=(11, 0, 0), file_labels=perm)
    _check(ref, RC.relabel_key("rmat_s12", 0, (11, 0, 0), 1, "file"), g, "file labels")
    g.free()
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
        g.free()
    g0.free()


def test_relabel_through_reference_api(golden, tmp_path):
    L = hostapi.load(False)
    z = golden("rmat_s10")
    ref = golden("reference_api")
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    for levels, dflag in RC.API_LEVELS:
        a = L.argumentsNew()
        keep = path.encode()
        a.contents.fnameb = keep
        a.contents.sort = 1
        a.contents.dflag = dflag
        a.contents.lmode, a.contents.lmode_l2, a.contents.lmode_l3 = levels
        g = L.graphCSRPreProcessingStep(a)
        gc = g.contents
        N, E = gc.num_vertices, gc.num_edges
        el = gc.sorted_edges_array.contents
        lab = np.ctypeslib.as_array(el.label_array, (N,)).copy()
        inv = np.ctypeslib.as_array(el.inverse_label_array, (N,))
        np.testing.assert_array_equal(inv[lab], np.arange(N))
        sides = {}
        for side, (v, e) in enumerate(((gc.vertices, gc.sorted_edges_array),
                                       (gc.inverse_vertices, gc.inverse_sorted_edges_array))):
            sides[side] = dict(out_degree=np.ctypeslib.as_array(v.contents.out_degree, (N,)).copy(),
                               edges_idx=np.ctypeslib.as_array(v.contents.edges_idx, (N,)).copy(),
                               src=np.ctypeslib.as_array(e.contents.edges_array_src, (E,)).copy(),
                               dst=np.ctypeslib.as_array(e.contents.edges_array_dest, (E,)).copy())
        _check_arrays(ref, RC.relabel_key("rmat_s10", dflag, levels, 1), lab, sides, f"levels {levels}")
        L.graphCSRFree(g)
        L.argumentsFree(a)
