"""The reference's data-driven push and pull-push PageRank on the GPU
(pageRankDataDriven{Push,PullPush}GraphCSR, pageRank.c:1872-2202; pushpull
codes 7 and 8) as frontier iterations with float atomics.

The reference updates ranks and residuals in place in vertex order, so its
output depends on the order (with several threads it changes from run to run:
1e-2 to 5e-2 L1 between 1, 2 and 8 threads at R-MAT scale 12).  What the
method determines is its limit as epsilon shrinks, which the reference's own
algorithm (the oracle, bit-exact with its one-thread outputs) reaches: the
push form's self-loop-free limit (oracle.dd_push_limit) and the pull-push
form's ordinary REF fixed point.  Bars, push form: within epsilon / (1 - d)
+ 1e-6 L1 of the limit (its residual is exactly the change of the rank still
to come: at the stop every residual is below epsilon, and (I - d P^T)^-1
grows an L1 norm by at most 1 / (1 - d)); and within twice that of the
reference's own output at epsilon 1e-4.  The pull-push form recomputes a rank
by a pull but pushes the residual, not the pull's change (pageRank.c:
2114-2143), so its stop rule bounds the pushed residuals only: a vertex
whose in-neighbours changed by more than their residuals can stay stale.
Which vertices that hits depends on the processing order (25 runs per graph
and epsilon on MI355X: 3e-8 to 6e-8 L1 from the fixed point in 49 of 50 runs
on fixture_v8192_e32768, 1.2e-5 in one); its bar is PULL_PUSH_BAR."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL, l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL] + ["ring_rmat_s10"]
MODES = ((gbx.PR_DD_PUSH, 7), (gbx.PR_DD_PULL_PUSH, 8))
D = 0.85
PULL_PUSH_BAR = 1e-4


def _bar(mode, eps):
    return eps / (1 - D) + 1e-6 if mode == gbx.PR_DD_PUSH else max(PULL_PUSH_BAR, eps / (1 - D) + 1e-6)


def _edges(golden, name):
    v = golden(f"variants_{name}")
    if "edges_src" in v:
        return v["edges_src"], v["edges_dst"]
    z = golden(name)
    return z["edges_src"], z["edges_dst"]


def _limits(g):
    f, inv = g.export(), g.export(inverse=True)
    rp = np.zeros(g.N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    lim7 = O.dd_push_limit(g.N, f["out_degree"], rp, inv["dst"])
    fix, _, _ = O.pr_f64(g.N, f["out_degree"], rp, inv["dst"], mode=0, iterations=3000, epsilon=1e-15)
    return {gbx.PR_DD_PUSH: lim7, gbx.PR_DD_PULL_PUSH: fix}


@pytest.mark.parametrize("name", GRAPHS)
def test_dd_push_reaches_the_methods_limits(gpu_ctx, golden, name):
    v = golden(f"variants_{name}")
    s, d = _edges(golden, name)
    print(f"[gpu] data-driven push {name}", flush=True)
    g = gpu_ctx.graph(s, d, sort_mode=1)
    lim = _limits(g)
    for mode, code in MODES:
        eps = 1e-7
        ranks, rep = g.pagerank(iterations=3000, epsilon=eps, mode=mode)
        e = l1(ranks, lim[mode])
        print(f"[gpu] {name} code {code}: {rep['iterations']} iterations, L1 vs limit {e:.2e}", flush=True)
        assert e <= _bar(mode, eps)
        eps = 1e-4
        ranks, rep = g.pagerank(iterations=3000, epsilon=eps, mode=mode)
        assert l1(ranks, lim[mode]) <= _bar(mode, eps)
        assert l1(ranks, v[f"pp{code}_eps_ranks"]) <= 2 * _bar(mode, eps)
        if code == 7:
            assert rep["converged"] == 1


def test_dd_push_rmat_s18(gpu_ctx):
    g = gpu_ctx.rmat_graph(18, 16, 20261203 + 18)
    print(f"[gpu] data-driven push R-MAT s18 N={g.N} E={g.E}", flush=True)
    lim = _limits(g)
    for mode, code in MODES:
        ranks, rep = g.pagerank(iterations=3000, epsilon=1e-7, mode=mode)
        e = l1(ranks, lim[mode])
        print(f"[gpu] s18 code {code}: {rep['iterations']} iterations, L1 vs limit {e:.2e}", flush=True)
        assert e <= _bar(mode, 1e-7)
        assert rep["rank_sum"] == pytest.approx(float(lim[mode].sum()), rel=1e-5)


def test_dd_push_self_loops_and_ring(gpu_ctx):
    """A ring with self loops on every other vertex: the push form drops the
    later self pushes (its limit is not the REF fixed point), the pull-push
    form reaches the REF fixed point."""
    n = 200
    v = np.arange(n, dtype=np.uint32)
    s = np.concatenate([v, v[::2]])
    d = np.concatenate([(v + 1) % n, v[::2]])
    g = gpu_ctx.graph(s, d, sort_mode=1)
    lim = _limits(g)
    assert l1(lim[gbx.PR_DD_PUSH], lim[gbx.PR_DD_PULL_PUSH]) > 1e-3
    for mode, _ in MODES:
        ranks, _ = g.pagerank(iterations=5000, epsilon=1e-8, mode=mode)
        assert l1(ranks, lim[mode]) <= _bar(mode, 1e-8)


def test_dd_push_graph_replay_and_errors(gpu_ctx, golden):
    s, d = _edges(golden, "rmat_s10")
    g = gpu_ctx.graph(s, d, sort_mode=1)
    lim = _limits(g)
    ses = g.session(iterations=400, epsilon=1e-7, stop=gbx.STOP_ACTIVE, mode=gbx.PR_DD_PUSH)
    for _ in range(3):
        ses.reset()
        ses.launch_graph(150)
        ses.launch_graph(250)
        r, rep = ses.read()
        assert rep["converged"] == 1
        assert l1(r, lim[gbx.PR_DD_PUSH]) <= 1e-7 / (1 - D) + 1e-6
    ses.free()
    with pytest.raises(gbx.GbxError, match="data-driven"):
        g.pagerank(iterations=5, mode=gbx.PR_DD_PUSH, method=gbx.METHOD_BINNED)
    with pytest.raises(gbx.GbxError, match="data-driven"):
        g.pagerank(iterations=5, mode=gbx.PR_DD_PULL_PUSH, precision=gbx.BF16)
    with pytest.raises(gbx.GbxError, match="teleport"):
        g.pagerank(iterations=5, mode=gbx.PR_DD_PUSH, teleport=np.full(g.N, 1.0 / g.N))
