"""Write tests/golden/reference_api.npz: outputs of the GraphBlox reference's
own code for the GPU tests that compare against its relabel pipeline, its
rank sort and its SpMV variants, so that those tests never load the
reference build.

Run in a container that has the reference sources (``/root/reference``):

    make -C oracle ref        # builds oracle/_ref/libgbxref_d<D>w<W>.so
    python tests/golden/make_reference_api_goldens.py

Stored (arrays as SHA-256 digests of their bytes, labels of the small graphs
in full):
  * relabel: graphCSRPreProcessingStep with relabel levels (graphCSR.c:
    268-290, reorderGraphProcess reorder.c:1556-1616) -- the label array and
    the forward and transpose out_degree / edges_idx / src / dst -- for every
    order on three graphs with and without -k, composed levels with sort
    modes 0 and 1, and a label file (order 11);
  * rank sort: radixSortEdgesByPageRank (reorder.c:165-221) of the seeded
    cases the rank-sort test draws (the ones inside the range C defines the
    float-to-u32 conversion for);
  * SpMV variants: SPMVGraphCSR (SPMV.c:89-590) codes 0-3, 3 passes, on
    R-MAT s12, unweighted (WEIGHTED=0 build) and weighted (WEIGHTED=1 build);
  * DIRECTED=0 builds (graphConfig.h:14 as shipped): on R-MAT s12, sort mode
    1, SPMVGraphCSR codes 0-3 (3 passes, unweighted and weighted) and the
    pageRankGraphCSR variants 0-5 and 9 after 20 iterations (epsilon 0): the
    pull forms over the forward rows (pageRank.c:269-272, SPMV.c:141-147),
    the push forms along the forward out-edges (in-edge sums).
The inputs are regenerated by the tests from the same goldens and seeds.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from oracle import oracle as O  # noqa: E402

import refapi_cases as RC  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def ref_relabel(path, dflag, levels, fnamel=None, sort=1):
    R = O.ref(1, 0)
    g = R.refd_build_relabel(path.encode(), sort, dflag, 0, 0, 1, *levels, fnamel.encode() if fnamel else None)
    N, E = R.refd_num_vertices(g), R.refd_num_edges(g)
    lab = np.zeros(N, np.uint32)
    R.refd_labels(g, lab.ctypes.data, None)
    arrays = {}
    for side in (0, 1):
        a = [np.zeros(N, np.uint32) for _ in range(3)] + [np.zeros(E, np.uint32) for _ in range(2)]
        R.refd_export(g, side, *[x.ctypes.data for x in a], None)
        arrays[side] = a
    R.refd_free(g)
    return lab, arrays


def put_relabel(out, key, lab, arrays, full_labels):
    out[f"{key}/labels_sha"] = np.array(sha(lab))
    if full_labels:
        out[f"{key}/labels"] = lab
    for side in (0, 1):
        od, _ind, idx, src, dst = arrays[side]
        for nm, a in (("out_degree", od), ("edges_idx", idx), ("src", src), ("dst", dst)):
            out[f"{key}/{side}/{nm}_sha"] = np.array(sha(a))


def ref_rank_sort(ranks, labels):
    R = O.ref(1, 0)
    R.radixSortEdgesByPageRank.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    R.radixSortEdgesByPageRank.restype = C.POINTER(C.c_uint32)
    r = np.ascontiguousarray(ranks, np.float32).copy()
    lab = np.ascontiguousarray(labels, np.uint32).copy()
    res = R.radixSortEdgesByPageRank(r.ctypes.data, lab.ctypes.data, r.size)
    return np.ctypeslib.as_array(res, (r.size,)).copy()


def main():
    out = {}
    with tempfile.TemporaryDirectory() as td:
        # relabel, one level, every order
        for name in RC.RELABEL_GRAPHS:
            z = load(name)
            path = os.path.join(td, f"{name}.bin")
            O.write_bin(path, z["edges_src"], z["edges_dst"])
            for dflag in (0, 1):
                for lm in RC.LMODES:
                    lab, arrays = ref_relabel(path, dflag, (lm, 0, 0))
                    put_relabel(out, RC.relabel_key(name, dflag, (lm, 0, 0), 1), lab, arrays, name != "rmat_s12")
            print("relabel", name, flush=True)
        z = load("rmat_s12")
        path = os.path.join(td, "rmat_s12.bin")
        O.write_bin(path, z["edges_src"], z["edges_dst"])
        for levels in RC.LEVELS:
            for sort in (0, 1):
                lab, arrays = ref_relabel(path, 0, levels, sort=sort)
                put_relabel(out, RC.relabel_key("rmat_s12", 0, levels, sort), lab, arrays, False)
        perm = RC.file_labels(int(z["m1_N"]))
        fl = os.path.join(td, "labels.txt")
        np.savetxt(fl, perm, fmt="%u")
        lab, arrays = ref_relabel(path, 0, (11, 0, 0), fnamel=fl)
        put_relabel(out, RC.relabel_key("rmat_s12", 0, (11, 0, 0), 1, "file"), lab, arrays, False)
        z10 = load("rmat_s10")
        p10 = os.path.join(td, "rmat_s10.bin")
        O.write_bin(p10, z10["edges_src"], z10["edges_dst"])
        for levels, dflag in RC.API_LEVELS:
            lab, arrays = ref_relabel(p10, dflag, levels)
            put_relabel(out, RC.relabel_key("rmat_s10", dflag, levels, 1), lab, arrays, True)
        print("relabel levels, file, API cases", flush=True)
        # rank sort
        for i, (what, ranks, labels, defined) in enumerate(RC.rank_sort_cases()):
            if defined:
                out[f"ranksort/{i}_sha"] = np.array(sha(ref_rank_sort(ranks, labels)))
        print("rank sort", flush=True)
        # SpMV variants through the reference's dispatch
        for weighted in (0, 1):
            R = O.ref(1, weighted)
            p = os.path.join(td, "g.wbin" if weighted else "g.bin")
            if weighted:
                O.write_bin(p, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
            else:
                O.write_bin(p, z["edges_src"], z["edges_dst"])
            g = R.refd_build(p.encode(), 1, 0, 0, weighted, 1)
            N = R.refd_num_vertices(g)
            for code in (0, 1, 2, 3):
                y, x = np.zeros(N, np.float32), np.zeros(N, np.float32)
                R.refd_spmv_variant(g, code, 3, 1, y, x)
                out[f"spmv/w{weighted}/code{code}"] = y
            R.refd_free(g)
        print("SpMV variants", flush=True)
        # the DIRECTED=0 builds: pulls over the forward CSR
        for weighted in (0, 1):
            R = O.ref(0, weighted)
            p = os.path.join(td, "g0.wbin" if weighted else "g0.bin")
            if weighted:
                O.write_bin(p, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
            else:
                O.write_bin(p, z["edges_src"], z["edges_dst"])
            g = R.refd_build(p.encode(), 1, 0, 0, weighted, 1)
            N = R.refd_num_vertices(g)
            for code in (0, 1, 2, 3):
                y, x = np.zeros(N, np.float32), np.zeros(N, np.float32)
                R.refd_spmv_variant(g, code, 3, 1, y, x)
                out[f"d0/spmv/w{weighted}/code{code}"] = y
            if not weighted:
                for code in (0, 1, 2, 3, 4, 5, 9):
                    r = np.zeros(N, np.float32)
                    it, err = C.c_uint32(), C.c_double()
                    R.refd_pagerank_variant(g, code, 20, 0.0, 1, r, C.byref(it), C.byref(err))
                    out[f"d0/pr/code{code}/r20"] = r
            R.refd_free(g)
        print("DIRECTED=0 builds", flush=True)
    dst = os.path.join(HERE, "reference_api.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes,", len(out), "entries")


if __name__ == "__main__":
    main()
