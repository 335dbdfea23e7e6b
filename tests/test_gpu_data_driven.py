"""The reference's data-driven pull PageRank on the GPU
(pageRankDataDrivenPullGraphCSR, pageRank.c:1711-1870; pushpull code 6):
worklist rows only, a row keeps its rank unless it changed by epsilon or
more, and the out-neighbours of updated rows form the next worklist.

Bars: the pull iteration's (row sums are FP32 sums in another order than the
reference's sequential one): <= 1e-5 L1 against the reference's own outputs.
With epsilon > 0 the stopping iteration is not compared: it is decided by the
last few rows whose change hovers at epsilon, and the summation order alone
moves it (the oracle's own algorithm stops at iteration 76, 82 or 90 on
fixture v8192_e32768 with sequential, reversed or FP64-rounded row sums, all
within 3.2e-6 L1 of each other), so the bar is convergence and the L1 at the
stop.  With epsilon 0 the data-driven pull is the plain pull (a row leaves the
worklist only when its in-neighbours stop changing), so it must be
bit-identical to the GPU's REF pull."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL, l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL] + ["ring_rmat_s10", "ring_64"]
DD = gbx.PR_DATA_DRIVEN
TOL = 1e-5


def _graph(ctx, golden, name):
    v = golden(f"variants_{name}")
    if "edges_src" in v:
        return ctx.graph(v["edges_src"], v["edges_dst"], sort_mode=1)
    z = golden(name)
    return ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)


@pytest.mark.parametrize("name", GRAPHS)
def test_data_driven_pull_vs_reference(gpu_ctx, golden, name):
    v = golden(f"variants_{name}")
    print(f"[gpu] data-driven pull {name}", flush=True)
    g = _graph(gpu_ctx, golden, name)
    ranks, rep = g.pagerank(iterations=20, epsilon=0.0, mode=DD)
    assert rep["executed"] == 20
    assert l1(ranks, v["pp6_r20"]) <= TOL
    assert rep["error_total"] == pytest.approx(float(v["pp6_err20"]), rel=1e-4, abs=1e-12)
    ref, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_PULL)
    np.testing.assert_array_equal(ranks, ref)
    ranks, rep = g.pagerank(iterations=100, epsilon=1e-4, mode=DD)
    it_ref = int(v["pp6_eps_iters"])
    print(f"[gpu] {name}: iterations {rep['iterations']} (reference {it_ref}), "
          f"L1 {l1(ranks, v['pp6_eps_ranks']):.2e}", flush=True)
    assert rep["converged"] == 1
    assert l1(ranks, v["pp6_eps_ranks"]) <= TOL


def test_data_driven_pull_rmat_s18_vs_oracle(gpu_ctx):
    g = gpu_ctx.rmat_graph(18, 16, 20261203 + 18)
    print(f"[gpu] data-driven pull R-MAT s18 N={g.N} E={g.E}", flush=True)
    f, inv = g.export(), g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    frp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(f["out_degree"], out=frp[1:])
    for eps, iters in ((0.0, 100), (1e-3, 200), (1e-4, 200)):
        ranks, rep = g.pagerank(iterations=iters, epsilon=eps, mode=DD)
        ro, it, _ = O.pr_pull_dd(g.N, f["out_degree"], rp, inv["dst"], frp, f["dst"], iters, eps)
        e = l1(ranks, ro)
        print(f"[gpu] s18 eps {eps}: iterations {rep['iterations']} (oracle {it}), L1 {e:.2e}", flush=True)
        assert e <= TOL
        if eps == 0.0:
            assert rep["iterations"] == it
        else:
            assert rep["converged"] == 1


def test_data_driven_pull_graph_replay_and_errors(gpu_ctx, golden):
    v = golden("variants_rmat_s10")
    g = _graph(gpu_ctx, golden, "rmat_s10")
    eager, _ = g.pagerank(iterations=30, epsilon=1e-5, mode=DD, stop=gbx.STOP_NONE)
    s = g.session(iterations=30, epsilon=1e-5, stop=gbx.STOP_NONE, mode=DD)
    for _ in range(3):
        s.reset()
        s.launch_graph(11)
        s.launch_graph(19)
        r, rep = s.read()
        assert rep["executed"] == 30
        np.testing.assert_array_equal(r, eager)
    s.free()
    assert l1(g.pagerank(iterations=20, epsilon=0.0, mode=DD)[0], v["pp6_r20"]) <= TOL
    with pytest.raises(gbx.GbxError, match="data-driven"):
        g.pagerank(iterations=5, mode=DD, method=gbx.METHOD_BINNED)
    with pytest.raises(gbx.GbxError, match="data-driven"):
        g.pagerank(iterations=5, mode=DD, precision=gbx.BF16)
    with pytest.raises(gbx.GbxError, match="teleport"):
        g.pagerank(iterations=5, mode=DD, teleport=np.full(g.N, 1.0 / g.N))
