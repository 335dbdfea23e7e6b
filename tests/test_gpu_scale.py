"""Parity at the workloads' own sizes, against the FP64 oracle and the port of
pageRankPullGraphCSR / SPMVPullGraphCSR (bit-exact with the reference's own
build on every graph the goldens hold, tests/test_oracle.py):

* R-MAT s16 (65,536 vertices): stochastic PageRank (uniform dangling mass),
  20 iterations, both methods, against FP64;
* R-MAT s22 (4,194,304 vertices): stochastic PageRank to L1 < 1e-6 (at most
  100 iterations) against FP64 with the stopping iteration within one and
  the ranks summing to 1; the uniform(0,1)-weighted SpMV (SPMV.c:167-194)
  against the port;
* R-MAT s24 (16,777,216 vertices): reference-semantics PageRank after 20
  iterations, both methods, against the port and FP64; the precision form
  (BF16 contributions) and FP32 to tolerance 1e-7 against FP64 (L1 and the
  largest relative error reported);
* R-MAT s26 (67,108,864 vertices, 1.06 billion edges): reference semantics
  after 20 iterations against the port and FP64, stochastic ranks summing
  to 1;
* 256 independent R-MAT graphs with 2^10..2^18 vertices (log-uniform) and
  edge factors 8..32 in one batch, every graph against FP64;
* 32 personalised vectors (seed sets of 1..1000 vertices) at s22, eight of
  them against FP64.
The oracle runs on the host's OpenMP threads (OMP_NUM_THREADS)."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 24  # the benchmark's graphs


def _oracle_of(g):
    f, i = g.export(False), g.export(True)
    mk = lambda c: O.OracleCSR(g.N, g.E, c["out_degree"], c["in_degree"], c["edges_idx"], c["src"], c["dst"], c["w"])
    fwd, inv = mk(f), mk(i)
    return fwd, inv, O.row_ptr_from(inv)


def _max_rel(r, r64):
    return float(np.max(np.abs(np.asarray(r, np.float64) - r64) / np.maximum(r64, 1e-300)))


@pytest.fixture(scope="module")
def s22(gpu_ctx):
    g = gpu_ctx.rmat_graph(22, 16, SEED)
    print(f"[gpu] R-MAT s22: N={g.N} E={g.E}", flush=True)
    yield g, _oracle_of(g)
    g.free()


@pytest.fixture(scope="module")
def s24(gpu_ctx):
    g = gpu_ctx.rmat_graph(24, 16, SEED)
    print(f"[gpu] R-MAT s24: N={g.N} E={g.E}", flush=True)
    yield g, _oracle_of(g)
    g.free()


@pytest.mark.parametrize("method", [gbx.METHOD_PULL, gbx.METHOD_BINNED])
def test_s16_stochastic_vs_fp64(gpu_ctx, method):
    g = gpu_ctx.rmat_graph(16, 16, SEED)
    fwd, inv, rp = _oracle_of(g)
    r, rep = g.pagerank(iterations=20, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE, method=method)
    r64, _, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=20)
    e = l1(r, r64)
    print(f"[gpu] s16 stochastic method {method}: L1 vs FP64 {e:.3e}, max rel {_max_rel(r, r64):.3e}", flush=True)
    assert e <= 1e-5
    assert abs(rep["rank_sum"] - 1.0) < 1e-5
    g.free()


@pytest.mark.parametrize("method", [gbx.METHOD_BINNED, gbx.METHOD_PULL])
def test_s22_stochastic_to_convergence_vs_fp64(s22, method):
    g, (fwd, inv, rp) = s22
    r, rep = g.pagerank(iterations=100, epsilon=1e-6, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, method=method)
    r64, it64, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, epsilon=1e-6)
    e = l1(r, r64)
    print(f"[gpu] s22 stochastic method {method}: {rep['executed']} iterations (FP64 {it64}), L1 vs FP64 {e:.3e}, "
          f"max rel {_max_rel(r, r64):.3e}", flush=True)
    assert rep["converged"] == 1 and abs(rep["executed"] - it64) <= 1
    assert e <= 1e-5
    assert abs(rep["rank_sum"] - 1.0) < 1e-5


def test_s22_weighted_spmv_vs_port(gpu_ctx):
    g = gpu_ctx.rmat_graph(22, 16, SEED, weight_kind=1, weight_seed=SEED, flags=gbx.GRAPH_WEIGHTED)
    fwd, inv, _ = _oracle_of(g)
    assert float(inv.w.min()) >= 0.0 and float(inv.w.max()) < 1.0 and abs(float(inv.w.mean()) - 0.5) < 1e-3
    x, y, sec = g.spmv(iterations=1, weighted=True)
    xo, yo = O.spmv_pull_f32(fwd, inv, iterations=1, weights=inv.w, nthreads=0)
    np.testing.assert_array_equal(x, xo)
    rel = l1(y, yo) / float(np.abs(yo.astype(np.float64)).sum())
    print(f"[gpu] s22 weighted SpMV: {sec * 1e3:.3f} ms, relative L1 vs port {rel:.3e}", flush=True)
    np.testing.assert_allclose(y, yo, rtol=2e-5, atol=1e-9)
    assert rel <= 1e-6
    g.free()


@pytest.mark.parametrize("method", [gbx.METHOD_BINNED, gbx.METHOD_PULL])
def test_s24_reference_semantics_vs_port_and_fp64(s24, method):
    g, (fwd, inv, rp) = s24
    r, rep = g.pagerank(iterations=20, epsilon=0.0, method=method)
    port = O.pr_pull_f32(fwd, inv, 20, 0.0, nthreads=0)["ranks"]
    r64, _, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=0, iterations=20)
    e_port, e64 = l1(r, port), l1(r, r64)
    print(f"[gpu] s24 REF 20 iterations method {method}: L1 vs port {e_port:.3e}, vs FP64 {e64:.3e}, "
          f"max rel vs FP64 {_max_rel(r, r64):.3e}; port vs FP64 {l1(port, r64):.3e}", flush=True)
    assert rep["executed"] == 20
    assert e_port <= 1e-5 and e64 <= 1e-5


def test_s24_precision_form_vs_fp64(s24):
    """BF16 contributions (binned: exact fixed-point row sums of the BF16
    values; FP64 reductions) and FP32, stochastic, to tolerance 1e-7.

    With BF16 contributions (and BF16 run sums) the iteration is a map on a
    finite set of values: once the change of the exact iteration falls below
    the rounding, it ends in a cycle (measured at R-MAT s24: period 11 with
    L1 changes of 1.4e-7..2.2e-7 for the binned method, period 8 with
    2.5e-8..1e-7 for the pull kernel), so the 1e-7 stop is reached only if a
    cycle dips below it.  The BF16 bar is its error against FP64 and a last
    change of at most 1e-6."""
    g, (fwd, inv, rp) = s24
    r64, it64, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=300, epsilon=1e-10)
    for prec, bar in ((gbx.F32, 1e-5), (gbx.BF16, 2e-3)):
        r, rep = g.pagerank(iterations=300, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, precision=prec)
        e = l1(r, r64)
        print(f"[gpu] s24 stochastic precision {prec}: {rep['executed']} iterations (converged "
              f"{rep['converged']}, last change {rep['error_total']:.3e}), L1 vs FP64 {e:.3e}, "
              f"max rel {_max_rel(r, r64):.3e} (FP64 to 1e-10: {it64} iterations)", flush=True)
        if prec == gbx.F32:
            assert rep["converged"] == 1
        else:
            assert rep["converged"] == 1 or rep["error_total"] <= 1e-6
        assert e <= bar
        assert abs(rep["rank_sum"] - 1.0) < 1e-5


def test_s26_binned_vs_port_and_fp64(gpu_ctx):
    """The largest graph of the suite: R-MAT s26 (67,108,864 vertices, 1.06
    billion edges), reference semantics with the binned iteration (the
    library's default at this size) after 20 iterations against the port and
    FP64, and stochastic ranks summing to 1."""
    g = gpu_ctx.rmat_graph(26, 16, 26)
    print(f"[gpu] R-MAT s26: N={g.N} E={g.E}", flush=True)
    try:
        r, rep = g.pagerank(iterations=20, epsilon=0.0)
        rs, reps = g.pagerank(iterations=20, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE)
        fwd, inv, rp = _oracle_of(g)
    finally:
        g.free()
    port = O.pr_pull_f32(fwd, inv, 20, 0.0, nthreads=0)["ranks"]
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=0, iterations=20)
    e_port, e64 = l1(r, port), l1(r, r64)
    print(f"[gpu] s26 REF 20 iterations: L1 vs port {e_port:.3e}, vs FP64 {e64:.3e}, "
          f"max rel vs FP64 {_max_rel(r, r64):.3e}", flush=True)
    assert rep["executed"] == 20
    assert e_port <= 1e-5 and e64 <= 1e-5
    assert abs(reps["rank_sum"] - 1.0) < 1e-5 and abs(float(np.sum(rs, dtype=np.float64)) - 1.0) < 1e-5


def test_batch_of_256_graphs_at_full_sizes(gpu_ctx):
    """256 graphs, vertex counts 2^10..2^18 (log-uniform), edge factors
    8..32 (uniform), one launch per iteration for all of them (captured as a
    HIP graph); every member against its own FP64 PageRank."""
    rng = np.random.default_rng(256)
    scales = rng.integers(10, 19, 256)
    efs = rng.integers(8, 33, 256)
    members = [gpu_ctx.rmat_graph(int(s), int(f), 5000 + k) for k, (s, f) in enumerate(zip(scales, efs))]
    b = gpu_ctx.graph_batch(members)
    off = b.batch_offsets()
    print(f"[gpu] batch of 256: N={b.N} E={b.E}, members {min(m.N for m in members)}..{max(m.N for m in members)} "
          f"vertices", flush=True)
    s = b.session(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
    s.reset()
    s.launch_graph(100)
    ranks, reps = s.read()
    s.free()
    worst = 0.0
    for k, m in enumerate(members):
        fwd, inv, rp = _oracle_of(m)
        r64, it64, _ = O.pr_f64(m.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, epsilon=1e-7)
        e = l1(ranks[off[k]:off[k + 1]], r64)
        worst = max(worst, e)
        assert reps[k]["converged"] == 1 and abs(reps[k]["executed"] - it64) <= 1, k
        assert e <= 1e-5, (k, e)
        assert abs(reps[k]["rank_sum"] - 1.0) < 1e-5, k
    print(f"[gpu] batch of 256: worst member L1 vs FP64 {worst:.3e}", flush=True)
    b.free()
    for m in members:
        m.free()


def test_32_personalised_vectors_at_s22(s22):
    g, (fwd, inv, rp) = s22
    rng = np.random.default_rng(SEED)
    sets = [rng.choice(g.N, int(rng.integers(1, 1001)), replace=False) for _ in range(32)]
    s = g.ppr_session(32, sets, iterations=20, epsilon=0.0, stop=gbx.STOP_NONE)
    s.reset()
    s.launch(20)
    ranks, reps = s.read()
    s.free()
    for q in range(0, 32, 4):
        t = np.zeros(g.N)
        t[sets[q]] = 1.0 / len(sets[q])
        r64, _, _ = O.pr_f64(g.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=20, teleport=t)
        e = l1(ranks[:, q], r64)
        print(f"[gpu] s22 personalised vector {q} ({len(sets[q])} seeds): L1 vs FP64 {e:.3e}", flush=True)
        assert e <= 1e-5
        assert abs(reps[q]["rank_sum"] - 1.0) < 1e-5
