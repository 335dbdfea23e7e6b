"""Write the golden vectors under tests/golden/ from the GraphBlox reference.

Run in a container that has the reference sources (``/root/reference``):

    make -C oracle ref        # builds oracle/_ref/libgbxref_d<D>w<W>.so
    python tests/golden/make_goldens.py

For every graph it stores the inputs and the reference's own outputs:
  * forward and transpose CSR arrays (graphCSRPreProcessingStep,
    01_host/src/structures/graphCSR.c:243-327, DIRECTED=1) for sort modes 1
    (radix src+dst) and 2 (count src), plus mode 1 with -k dedupe;
  * FP32 ranks of pageRankPullGraphCSR (pageRank.c:245-368, DIRECTED=1) after
    1 and 20 iterations with epsilon 0, and the stopping iteration, error and
    ranks with epsilon 1e-4 (up to 100 iterations);
  * DIRECTED=0 ranks after 20 iterations (pull over the forward CSR);
  * SPMVPullGraphCSR (SPMV.c:117-215) x and y after 3 iterations;
  * for graphs with float weights, the WEIGHTED=1 build's weighted SpMV y.
Graphs: the five shipped CSR fixtures (03_data/TEST, whose text dumps are also
stored), and seeded R-MAT graphs (scale 8, 10, 12, edge factor 16) generated
by the oracle's generator.  For R-MAT scale 16 (the 65,536-vertex parity
case) only digests and the 20-iteration ranks are stored.
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
from oracle import oracle as O  # noqa: E402

REF_TEST = "/root/reference/03_data/TEST"
FIXTURES = ["v500_e500", "v1024_e4096", "v8192_e32768", "v1000_e10000", "v529_e3500"]
RMAT_SEED = 20261203


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_csr(L, path, mode, dflag, weighted=0):
    g = L.refd_build(path.encode(), mode, dflag, 0, weighted, 4)
    N, E = L.refd_num_vertices(g), L.refd_num_edges(g)
    out = {}
    for inv in (0, 1):
        a = dict(out_degree=np.zeros(N, np.uint32), in_degree=np.zeros(N, np.uint32),
                 edges_idx=np.zeros(N, np.uint32), src=np.zeros(E, np.uint32), dst=np.zeros(E, np.uint32),
                 w=np.zeros(E, np.float32))
        L.refd_export(g, inv, *(a[k].ctypes.data for k in ("out_degree", "in_degree", "edges_idx", "src", "dst", "w")))
        out[inv] = a
    return g, N, E, out


def ref_pr(L, g, N, iters, eps):
    r = np.zeros(N, np.float32)
    it, err, t = C.c_uint32(), C.c_double(), C.c_double()
    L.refd_pagerank(g, iters, eps, 4, r, C.byref(it), C.byref(err), C.byref(t))
    return r, it.value, err.value


def ref_spmv(L, g, N, iters):
    y = np.zeros(N, np.float32)
    x = np.zeros(N, np.float32)
    t = C.c_double()
    L.refd_spmv(g, iters, 4, y, x, C.byref(t))
    return x, y


def golden_for(path: str, name: str, weighted_path: str | None = None, full=True):
    d1 = O.ref(1, 0)
    d0 = O.ref(0, 0)
    out = {}
    src, dst, _, N_in = O.read_bin(path)
    out["edges_src"], out["edges_dst"] = src, dst
    for mode, dflag, tag in ((1, 0, "m1"), (2, 0, "m2"), (1, 1, "k")):
        g, N, E, csr = ref_csr(d1, path, mode, dflag)
        out[f"{tag}_N"], out[f"{tag}_E"] = N, E
        for inv, side in ((0, "fwd"), (1, "inv")):
            for k in ("out_degree", "in_degree", "edges_idx", "src", "dst"):
                out[f"{tag}_{side}_{k}"] = csr[inv][k]
        if tag == "m1":
            r1, _, _ = ref_pr(d1, g, N, 1, 0.0)
            r20, it20, err20 = ref_pr(d1, g, N, 20, 0.0)
            re, ite, erre = ref_pr(d1, g, N, 100, 1e-4)
            out.update(pr_r1=r1, pr_r20=r20, pr_err20=err20, pr_eps_ranks=re, pr_eps_iters=ite, pr_eps_err=erre)
            x, y = ref_spmv(d1, g, N, 3)
            out.update(spmv_x=x, spmv_y3=y)
        d1.refd_free(g)
    g0, N0, _, _ = ref_csr(d0, path, 1, 0)
    r0, _, _ = ref_pr(d0, g0, N0, 20, 0.0)
    out["pr_fwd_r20"] = r0
    d0.refd_free(g0)
    if weighted_path:
        d1w = O.ref(1, 1)
        ws, wd, ww, _ = O.read_bin(weighted_path, weighted=True)
        out.update(wedges_src=ws, wedges_dst=wd, wedges_w=ww)
        for mode, tag in ((1, "wm1"), (2, "wm2")):
            g, N, E, csr = ref_csr(d1w, weighted_path, mode, 0, weighted=1)
            for inv, side in ((0, "fwd"), (1, "inv")):
                out[f"{tag}_{side}_w"] = csr[inv]["w"]
                out[f"{tag}_{side}_dst"] = csr[inv]["dst"]
            if tag == "wm1":
                x, y = ref_spmv(d1w, g, N, 1)
                out.update(wspmv_x=x, wspmv_y1=y)
            d1w.refd_free(g)
    if not full:
        small = {k: v for k, v in out.items() if not k.startswith(("edges_", "m1_", "m2_", "k_", "wedges_", "wm"))}
        small["edges_sha"] = sha(np.stack([src, dst]))
        for tag in ("m1",):
            for side in ("fwd", "inv"):
                for k in ("out_degree", "edges_idx", "src", "dst"):
                    small[f"{tag}_{side}_{k}_sha"] = sha(out[f"{tag}_{side}_{k}"])
            small[f"{tag}_N"], small[f"{tag}_E"] = out[f"{tag}_N"], out[f"{tag}_E"]
        for k in ("pr_r1", "pr_eps_ranks", "spmv_x", "spmv_y3", "pr_fwd_r20"):
            small.pop(k, None)
        out = small
    dest = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(dest, **{k: np.asarray(v) for k, v in out.items()})
    print("wrote", dest, os.path.getsize(dest), "bytes")


def main():
    if not (O.ref_available(1, 0) and O.ref_available(0, 0) and O.ref_available(1, 1)):
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    for fx in FIXTURES:
        path = os.path.join(REF_TEST, fx, "graph.bin")
        golden_for(path, f"fixture_{fx}")
        # the reference's own text dumps of the forward CSR
        t = O.read_csr_text_fixture(path)
        np.savez_compressed(os.path.join(HERE, f"textdump_{fx}.npz"),
                            **{k: np.asarray(v) for k, v in t.items() if v is not None})
    rng = np.random.default_rng(RMAT_SEED)
    with tempfile.TemporaryDirectory() as tmp:
        for scale in (8, 10, 12, 16):
            s, d, N = O.rmat_graph(scale, 16, RMAT_SEED + scale)
            p = os.path.join(tmp, f"rmat{scale}.bin")
            O.write_bin(p, s, d)
            wp = None
            if scale <= 12:
                w = rng.random(s.size, dtype=np.float32)
                wp = os.path.join(tmp, f"rmat{scale}.wbin")
                O.write_bin(wp, s, d, w)
            golden_for(p, f"rmat_s{scale}", weighted_path=wp, full=scale <= 12)


if __name__ == "__main__":
    main()
