"""The reference's SpMV variants on the GPU (SPMVGraphCSR, SPMV.c:89-590):
the fixed-point pull (2) and push (3) forms bit-exact with the oracle (itself
bit-exact with the reference) and with the reference's own outputs through
the adapter; the float push (1) sums the pull's terms (within the pull's
tolerance)."""
import ctypes as C

import numpy as np
import pytest

import graphblox_amd as gbx
import hostapi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _inv_rows(g):
    inv = g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    return inv, rp, g.export()["out_degree"]


@pytest.mark.parametrize("name", ["rmat_s10", "rmat_s12", "fixture_v8192_e32768"])
def test_spmv_fixed_bit_exact(gpu_ctx, golden, name):
    z = golden(name)
    print(f"[gpu] fixed SpMV {name}", flush=True)
    cases = [(False, z["edges_src"], z["edges_dst"], None)]
    if "wedges_src" in z:
        cases.append((True, z["wedges_src"], z["wedges_dst"], z["wedges_w"]))
    for weighted, s, d, w in cases:
        flags = gbx.GRAPH_WEIGHTED if weighted else 0
        g = gpu_ctx.graph(s, d, w, sort_mode=1, flags=flags)
        inv, rp, fod = _inv_rows(g)
        for bits in (32, 64):
            for it in (1, 3):
                x, y, _ = g.spmv_fixed(bits, iterations=it, weighted=weighted)
                xo, yo = O.spmv_fixed(g.N, fod, rp, inv["dst"], inv["w"] if weighted else None, bits, it)
                np.testing.assert_array_equal(x, xo)
                np.testing.assert_array_equal(y, yo, err_msg=f"bits {bits} weighted {weighted} iterations {it}")


def test_spmv_fixed_rmat_s18_hubs(gpu_ctx):
    """Rows of tens of thousands of edges spread over many waves (atomic
    combines of partial row sums)."""
    rng = np.random.default_rng(7)
    g0 = gpu_ctx.rmat_graph(18, 16, 20261203 + 18)
    f = g0.export()
    w = rng.random(f["src"].size).astype(np.float32)
    g = gpu_ctx.graph(f["src"], f["dst"], w, sort_mode=1, flags=gbx.GRAPH_WEIGHTED)
    inv, rp, fod = _inv_rows(g)
    assert inv["out_degree"].max() > 10000
    for bits in (32, 64):
        x, y, _ = g.spmv_fixed(bits, iterations=2, weighted=True)
        _, yo = O.spmv_fixed(g.N, fod, rp, inv["dst"], inv["w"], bits, 2)
        np.testing.assert_array_equal(y, yo)


@pytest.mark.parametrize("weighted", [False, True])
def test_spmv_variants_through_reference_api(golden, tmp_path, weighted):
    """SPMVGraphCSR codes 0-3 through the adapter against the reference's own
    SPMVGraphCSR outputs on the same file (recorded by
    tests/golden/make_reference_api_goldens.py)."""
    L = hostapi.load(weighted)
    L.SPMVGraphCSR.restype = C.POINTER(hostapi.SPMVStats)
    z = golden("rmat_s12")
    ref = golden("reference_api")
    path = str(tmp_path / ("g.wbin" if weighted else "g.bin"))
    if weighted:
        O.write_bin(path, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
    else:
        O.write_bin(path, z["edges_src"], z["edges_dst"])
    a = L.argumentsNew()
    keep = path.encode()
    a.contents.fnameb = keep
    a.contents.sort = 1
    a.contents.weighted = int(weighted)
    a.contents.iterations = 3
    g = L.graphCSRPreProcessingStep(a)
    N = g.contents.num_vertices
    for code in (0, 1, 2, 3):
        a.contents.pushpull = code
        st = L.SPMVGraphCSR(a, g)
        assert st, code
        y = np.ctypeslib.as_array(st.contents.vector_output, (N,)).copy()
        yr = ref[f"spmv/w{int(weighted)}/code{code}"]
        if code >= 2:
            np.testing.assert_array_equal(y, yr, err_msg=f"code {code}")
        else:
            assert np.abs(y.astype(np.float64) - yr).sum() <= 1e-6 * np.abs(yr.astype(np.float64)).sum()
        L.freeSPMVStats(st)
    L.graphCSRFree(g)
    L.argumentsFree(a)
