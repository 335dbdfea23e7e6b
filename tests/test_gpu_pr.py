"""GPU pull PageRank / SpMV parity: against the reference's FP32 outputs
(golden vectors, <= 1e-5 L1) and against the FP64 oracle."""
import numpy as np
import pytest

import graphblox_amd as gbx
from conftest import FIXTURES, RMAT_SMALL, l1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL]
TOL_F32 = 1e-5  # L1 on normalised ranks vs the reference / FP64 (the FP32 parity bar)


def _graph(ctx, z, **kw):
    return ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1, **kw)


def _oracle_csr(z):
    return O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)


@pytest.mark.parametrize("name", GRAPHS)
def test_pr_ref_matches_reference(gpu_ctx, golden, name):
    z = golden(name)
    print(f"[gpu] pr REF {name}", flush=True)
    g = _graph(gpu_ctx, z)
    for iters, key in ((1, "pr_r1"), (20, "pr_r20")):
        ranks, rep = g.pagerank(iterations=iters, epsilon=0.0)
        assert rep["iterations"] == iters and rep["executed"] == iters
        assert l1(ranks, z[key]) <= TOL_F32
        assert np.abs(ranks - z[key]).max() <= 1e-6 * max(1.0, float(z[key].max()) * g.N)
    assert rep["error_total"] == pytest.approx(float(z["pr_err20"]), rel=1e-4)
    # error against the FP64 oracle, reported and bounded
    fwd, inv = _oracle_csr(z)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, O.row_ptr_from(inv), inv.dst, mode=0, iterations=20)
    err64 = l1(ranks, r64)
    print(f"[gpu] {name}: L1 vs reference {l1(ranks, z['pr_r20']):.3e}, vs FP64 {err64:.3e}")
    assert err64 <= TOL_F32


@pytest.mark.parametrize("name", GRAPHS)
def test_pr_ref_epsilon_stop_matches_reference(gpu_ctx, golden, name):
    z = golden(name)
    g = _graph(gpu_ctx, z)
    ranks, rep = g.pagerank(iterations=100, epsilon=1e-4)
    assert abs(rep["iterations"] - int(z["pr_eps_iters"])) <= 1
    assert rep["converged"] == 1
    if rep["iterations"] == int(z["pr_eps_iters"]):
        assert l1(ranks, z["pr_eps_ranks"]) <= TOL_F32


@pytest.mark.parametrize("name", GRAPHS)
def test_pr_forward_direction_matches_reference(gpu_ctx, golden, name):
    """DIRECTED=0 build: the pull walks the forward CSR (pageRank.c:269-272)."""
    z = golden(name)
    g = _graph(gpu_ctx, z)
    ranks, _ = g.pagerank(iterations=20, epsilon=0.0, direction=gbx.PULL_FWD)
    assert l1(ranks, z["pr_fwd_r20"]) <= TOL_F32


@pytest.mark.parametrize("name", GRAPHS)
def test_spmv_matches_reference(gpu_ctx, golden, name):
    z = golden(name)
    g = _graph(gpu_ctx, z)
    x, y, _ = g.spmv(iterations=3)
    np.testing.assert_array_equal(x, z["spmv_x"])
    # rows summed by one lane are bit-exact; wider rows are reassociated
    np.testing.assert_allclose(y, z["spmv_y3"], rtol=2e-5, atol=1e-9)
    assert l1(y, z["spmv_y3"]) <= 1e-6 * float(np.abs(z["spmv_y3"]).sum())


@pytest.mark.parametrize("s", RMAT_SMALL)
def test_weighted_spmv_matches_reference(gpu_ctx, golden, s):
    z = golden(f"rmat_s{s}")
    g = gpu_ctx.graph(z["wedges_src"], z["wedges_dst"], z["wedges_w"], sort_mode=1, flags=gbx.GRAPH_WEIGHTED)
    x, y, _ = g.spmv(iterations=1, weighted=True)
    np.testing.assert_allclose(y, z["wspmv_y1"], rtol=2e-5, atol=1e-9)
    assert l1(y, z["wspmv_y1"]) <= 1e-6 * float(np.abs(z["wspmv_y1"]).sum())


@pytest.mark.parametrize("name", ["rmat_s10", "rmat_s12", "fixture_v8192_e32768"])
def test_pr_stochastic_matches_fp64_oracle(gpu_ctx, golden, name):
    z = golden(name)
    g = _graph(gpu_ctx, z)
    fwd, inv = _oracle_csr(z)
    rp = O.row_ptr_from(inv)
    ranks, rep = g.pagerank(iterations=100, epsilon=1e-6, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
    r64, it64, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=100, epsilon=1e-6)
    assert rep["converged"] == 1 and abs(rep["executed"] - it64) <= 1
    assert abs(rep["rank_sum"] - 1.0) < 1e-5
    exact = O.dense_stochastic(fwd.N, fwd.out_degree, rp, inv.dst) if fwd.N <= 2048 else None
    e = l1(ranks, r64)
    print(f"[gpu] stochastic {name}: L1 vs FP64 power iteration {e:.3e}")
    assert e <= TOL_F32
    if exact is not None:
        assert l1(ranks, exact) <= 1e-5


def test_pr_stochastic_personalised(gpu_ctx, golden):
    z = golden("rmat_s12")
    g = _graph(gpu_ctx, z)
    fwd, inv = _oracle_csr(z)
    rng = np.random.default_rng(11)
    seeds = rng.choice(fwd.N, 50, replace=False)
    t = np.zeros(fwd.N)
    t[seeds] = 1.0 / seeds.size
    ranks, rep = g.pagerank(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1, teleport=t)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, O.row_ptr_from(inv), inv.dst, mode=1, iterations=200,
                         epsilon=1e-7, teleport=t)
    assert l1(ranks, r64) <= TOL_F32


def test_pr_session_matches_one_shot(gpu_ctx, golden):
    z = golden("rmat_s12")
    g = _graph(gpu_ctx, z)
    r_one, _ = g.pagerank(iterations=20, epsilon=0.0)
    s = g.session(iterations=20, epsilon=0.0, stop=gbx.STOP_NONE)
    s.reset()
    s.launch(7)
    ms = s.launch_timed(13)
    r, rep = s.read()
    assert ms > 0 and rep["executed"] == 20
    np.testing.assert_array_equal(r, r_one)
    info = s.plan_info()
    assert sum(c["rows"] for c in info) == g.N
    assert sum(c["edges"] for c in info) == g.E
    # deterministic: same result again after reset
    s.reset()
    s.launch(20)
    r2, _ = s.read()
    np.testing.assert_array_equal(r2, r_one)


def test_pr_ref_s16_matches_reference(gpu_ctx, golden):
    """65,536-vertex R-MAT parity case, 20 iterations, FP32 vs the reference."""
    z = golden("rmat_s16")
    g = gpu_ctx.rmat_graph(16, 16, 20261203 + 16)
    ranks, rep = g.pagerank(iterations=20, epsilon=0.0)
    e = l1(ranks, z["pr_r20"])
    print(f"[gpu] s16 L1 vs reference {e:.3e}")
    assert e <= TOL_F32


def test_pr_scale_20_against_oracle_f32(gpu_ctx):
    """Hub chunks and every bin exercised at scale 20; compared with the
    oracle's FP32 restatement (itself bit-exact with the reference)."""
    print("[gpu] rmat s20", flush=True)
    g = gpu_ctx.rmat_graph(20, 16, 2020)
    fwd_e, inv_e = g.export(False), g.export(True)
    fwd = O.OracleCSR(g.N, g.E, fwd_e["out_degree"], fwd_e["in_degree"], fwd_e["edges_idx"], fwd_e["src"],
                      fwd_e["dst"], fwd_e["w"])
    inv = O.OracleCSR(g.N, g.E, inv_e["out_degree"], inv_e["in_degree"], inv_e["edges_idx"], inv_e["src"],
                      inv_e["dst"], inv_e["w"])
    ranks, rep = g.pagerank(iterations=20, epsilon=0.0)
    ref = O.pr_pull_f32(fwd, inv, 20, 0.0, nthreads=8)
    s = g.session(iterations=20, epsilon=0.0)
    info = s.plan_info()
    print("[gpu] s20 plan:", info)
    assert info[0]["rows"] > 0  # hub rows present
    e = l1(ranks, ref["ranks"])
    print(f"[gpu] s20 L1 vs oracle f32 {e:.3e}")
    assert e <= TOL_F32


@pytest.mark.parametrize("lds_bytes", ["0", "4096", "65536"])
def test_lds_hot_prefix_is_bit_identical(gpu_ctx, golden, tuning, lds_bytes):
    """The LDS copy of the hot contributions holds the same values, so the
    ranks must not change in any bit."""
    z = golden("rmat_s12")
    g = _graph(gpu_ctx, z)
    tuning(lds_bytes=0)
    r0, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_PULL)
    tuning(lds_bytes=int(lds_bytes))
    r1, _ = g.pagerank(iterations=20, epsilon=0.0, method=gbx.METHOD_PULL)
    np.testing.assert_array_equal(r0, r1)
    r2, _ = g.pagerank(iterations=20, epsilon=0.0, precision=gbx.BF16, method=gbx.METHOD_PULL)
    tuning(lds_bytes=0)
    r3, _ = g.pagerank(iterations=20, epsilon=0.0, precision=gbx.BF16, method=gbx.METHOD_PULL)
    np.testing.assert_array_equal(r2, r3)


@pytest.mark.parametrize("name", ["rmat_s12", "fixture_v8192_e32768", "fixture_v1000_e10000"])
def test_bf16_gather_error_bound(gpu_ctx, golden, name):
    """BF16 contribution gathers with Kahan FP32 row sums: L1 <= 2e-3 vs FP64."""
    z = golden(name)
    g = _graph(gpu_ctx, z)
    fwd, inv = _oracle_csr(z)
    rp = O.row_ptr_from(inv)
    rs, rep = g.pagerank(iterations=200, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1,
                         precision=gbx.BF16)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=200, epsilon=1e-9)
    e = l1(rs, r64)
    rel = float(np.max(np.abs(rs - r64) / r64))
    print(f"[gpu] bf16 stochastic {name}: L1 vs FP64 {e:.3e}, max rel {rel:.3e}, iters {rep['executed']}")
    assert e <= 2e-3
    assert abs(rep["rank_sum"] - 1.0) < 1e-5
    rr, _ = g.pagerank(iterations=20, epsilon=0.0, precision=gbx.BF16)
    assert l1(rr, z["pr_r20"]) <= 2e-3


def test_bf16_scale_20_error(gpu_ctx):
    g = gpu_ctx.rmat_graph(20, 16, 2020)
    r32, _ = g.pagerank(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
    r16, _ = g.pagerank(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1,
                        precision=gbx.BF16)
    e = l1(r16, r32)
    print(f"[gpu] s20 bf16 vs f32 L1 {e:.3e}")
    assert e <= 2e-3
