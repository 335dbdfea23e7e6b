"""Write tests/golden/adjoint_fd_s<scale>.npz: FP64 central finite
differences of the differentiable weighted PageRank loss.

Workload (oracle.weighted_rmat_workload): R-MAT scale 20, edge factor 16,
seeded, ids permuted, duplicates removed; FP32 edge weights lognormal(0, 1);
loss L = sum_v g[v] p[v] with g ~ N(0, 1); p the stochastic fixed point with
row-normalised weights, damping 0.85, uniform teleport.  No reference
counterpart: the values come from the oracle's forward solve only
(orc_pr_loss_diff, extended precision), never from an adjoint, so they are an
independent check of every adjoint implementation.

Stored: 64 edges (u, v) drawn uniformly from the edges whose source has at
least two out-edges (a single out-edge has P[u, v] = 1 for every weight, so
its gradient is identically zero), dL/dW at each by central differences with
relative step h, and the directional derivative sum_e dL/dW[e] xi[e] w[e]
along xi ~ N(0, 1) in edge-list order (a dot-product test of the whole
gradient).  Run: python tests/golden/make_adjoint_fd.py [scale]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

SEED = 20261203
H = 2e-3
ITERS = 100  # iterations of each perturbed solve from the converged p (0.85^100 < 1e-7)


def main(scale: int = 20) -> None:
    t0 = time.time()
    s, d, w, N, g = O.weighted_rmat_workload(scale, 16, SEED)
    fwd, inv = O.build_graph(s, d, w, N, mode=1, move_weights=True)
    rp = O.row_ptr_from(inv)
    wr = inv.w.astype(np.float64)
    p, _, _, _ = O.adjoint_f64(N, rp, inv.dst, wr, g, iterations=2000, tol=1e-15)
    rng = np.random.default_rng(SEED + 300 + scale)
    src_of = inv.dst
    eligible = np.flatnonzero(fwd.out_degree[src_of] >= 2)
    edges = np.sort(rng.choice(eligible, 64, replace=False))
    row = np.searchsorted(rp, edges, side="right") - 1
    fd = O.fd_edge_grads(N, rp, inv.dst, wr, g, edges, p, h=H, iterations=ITERS)
    print(f"64 edges {time.time() - t0:.0f} s", flush=True)
    # directional derivative along xi (edge-list order -> in-CSR order by key)
    xi_list = np.random.default_rng(SEED + 400 + scale).standard_normal(s.size)
    key_list = s.astype(np.uint64) * np.uint64(N) + d.astype(np.uint64)
    key_inv = inv.dst.astype(np.uint64) * np.uint64(N) + inv.src.astype(np.uint64)
    order = np.argsort(key_list)
    xi = xi_list[order][np.searchsorted(key_list[order], key_inv)]
    wp = wr * (1.0 + H * xi)
    wm = wr * (1.0 - H * xi)
    dir_fd = O.loss_diff(N, rp, inv.dst, wp, wm, g, p0=p, iterations=ITERS) / (2.0 * H)
    np.savez_compressed(os.path.join(HERE, f"adjoint_fd_s{scale}.npz"), scale=scale, edge_factor=16, seed=SEED,
                        N=N, E=s.size, h=H, iters=ITERS, u=inv.dst[edges], v=row.astype(np.uint32), fd=fd,
                        dir_fd=dir_fd, loss=float(np.dot(g, p)))
    print(f"done {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 20)
