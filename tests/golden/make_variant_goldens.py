"""Write tests/golden/variants_<graph>.npz: the GraphBlox reference's own
PageRank variants (pageRankGraphCSR, 01_host/src/algorithms/openmp/
pageRank.c:177-242, DIRECTED=1) on the golden graphs.

For each variant code c (arguments->pushpull) it stores the FP32 ranks
after 20 iterations with epsilon 0 (`pp<c>_r20`, `pp<c>_err20`) and, with
epsilon 1e-4 and up to 100 iterations, the ranks, the stopping iteration
and the error (`pp<c>_eps_ranks`, `pp<c>_eps_iters`, `pp<c>_eps_err`).
Codes: 1 push, 2 pull fixed point 64-bit, 3 push fixed point, 4 pull
quantised 32-bit, 5 push quantised, 6 data-driven pull, 7 data-driven push,
8 data-driven pull-push, 9/10/11 pull fixed point 32/16/8-bit, 12/13 pull
quantised 16/8-bit.  One OpenMP thread, so the float-atomic push variants
and the in-place data-driven push variants are deterministic.  The graphs are those of make_goldens.py (the
shipped fixtures and the seeded R-MAT scale 8-12 graphs), plus two graphs
without dangling vertices, whose edges are stored with them: `ring_rmat_s10`
(the R-MAT scale 10 graph and the ring v -> v+1 mod N: the quantised
variants see min > 0) and `ring_64` (a 64-vertex ring: every contribution
equal, min == max).

    make -C oracle ref && python tests/golden/make_variant_goldens.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

REF_TEST = "/root/reference/03_data/TEST"
FIXTURES = ["v500_e500", "v1024_e4096", "v8192_e32768", "v1000_e10000", "v529_e3500"]
RMAT_SEED = 20261203
CODES = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13)


def variants_for(path: str, name: str, edges=None):
    L = O.ref(1, 0)
    g = L.refd_build(path.encode(), 1, 0, 0, 0, 1)
    N = L.refd_num_vertices(g)
    out = {}
    for c in CODES:
        for iters, eps, tag in ((20, 0.0, "r20"), (100, 1e-4, "eps")):
            r = np.zeros(N, np.float32)
            it, err = C.c_uint32(), C.c_double()
            L.refd_pagerank_variant(g, c, iters, eps, 1, r, C.byref(it), C.byref(err))
            if tag == "r20":
                out[f"pp{c}_r20"], out[f"pp{c}_err20"] = r, err.value
            else:
                out[f"pp{c}_eps_ranks"], out[f"pp{c}_eps_iters"], out[f"pp{c}_eps_err"] = r, it.value, err.value
    L.refd_free(g)
    if edges is not None:
        out["edges_src"], out["edges_dst"] = edges
    dest = os.path.join(HERE, f"variants_{name}.npz")
    np.savez_compressed(dest, **{k: np.asarray(v) for k, v in out.items()})
    print("wrote", dest, os.path.getsize(dest), "bytes")


def main(only=None):
    """Writes every file, or only the graphs named on the command line."""
    if not O.ref_available(1, 0):
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    want = (lambda name: not only or name in only)
    for fx in FIXTURES:
        if want(f"fixture_{fx}"):
            variants_for(os.path.join(REF_TEST, fx, "graph.bin"), f"fixture_{fx}")
    with tempfile.TemporaryDirectory() as tmp:
        for scale in (8, 10, 12):
            s, d, N = O.rmat_graph(scale, 16, RMAT_SEED + scale)
            p = os.path.join(tmp, f"rmat{scale}.bin")
            O.write_bin(p, s, d)
            if want(f"rmat_s{scale}"):
                variants_for(p, f"rmat_s{scale}")
        s, d, N = O.rmat_graph(10, 16, RMAT_SEED + 10)
        v = np.arange(N, dtype=np.uint32)
        rs, rd, _ = O.dedupe_keep_first(np.concatenate([s, v]), np.concatenate([d, (v + 1) % N]))
        rings = {"ring_rmat_s10": (rs, rd),
                 "ring_64": (np.arange(64, dtype=np.uint32), (np.arange(64, dtype=np.uint32) + 1) % 64)}
        for name, (rs, rd) in rings.items():
            if not want(name):
                continue
            p = os.path.join(tmp, f"{name}.bin")
            O.write_bin(p, rs, rd)
            variants_for(p, name, (rs.astype(np.uint32), rd.astype(np.uint32)))


if __name__ == "__main__":
    main(set(sys.argv[1:]))
