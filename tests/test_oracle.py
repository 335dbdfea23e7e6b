"""CPU tests: pin the oracle against the reference's golden vectors and against
closed forms, and check the native libraries load and export their ABI."""
import os
import re

import numpy as np
import pytest

from conftest import FIXTURES, GOLDEN, RMAT_SMALL, ROOT, l1
from oracle import oracle as O

GRAPHS = [f"fixture_{f}" for f in FIXTURES] + [f"rmat_s{s}" for s in RMAT_SMALL]


# ---------------------------------------------------------------- CSR build

@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("tag,mode,dedupe", [("m1", 1, False), ("m2", 2, False), ("k", 1, True)])
def test_oracle_csr_matches_reference(golden, name, tag, mode, dedupe):
    z = golden(name)
    src, dst = z["edges_src"], z["edges_dst"]
    N = int(z[f"{tag}_N"])
    fwd, inv = O.build_graph(src, dst, None, N, mode=mode, dedupe=dedupe)
    assert fwd.E == int(z[f"{tag}_E"])
    for side, c in (("fwd", fwd), ("inv", inv)):
        for k in ("out_degree", "in_degree", "edges_idx", "src", "dst"):
            np.testing.assert_array_equal(getattr(c, k), z[f"{tag}_{side}_{k}"], err_msg=f"{side}.{k}")


@pytest.mark.parametrize("fx", FIXTURES)
def test_oracle_csr_matches_shipped_text_dumps(golden, fx):
    """The reference's own 03_data/TEST CSR dumps (sort src+dst, no dedupe)."""
    t = golden(f"textdump_{fx}")
    z = golden(f"fixture_{fx}")
    fwd, _ = O.build_graph(z["edges_src"], z["edges_dst"], None, int(t["N"]), mode=3)
    np.testing.assert_array_equal(fwd.out_degree, t["out_degree"])
    np.testing.assert_array_equal(fwd.edges_idx, t["edges_idx"])
    np.testing.assert_array_equal(fwd.src, t["src"])
    np.testing.assert_array_equal(fwd.dst, t["dst"])
    assert int(t["E"]) == fwd.E


def test_edges_idx_zero_for_empty_rows(golden):
    z = golden("fixture_v500_e500")
    od, idx = z["m1_fwd_out_degree"], z["m1_fwd_edges_idx"]
    assert (od == 0).sum() > 100
    assert np.all(idx[od == 0] == 0)


def test_count_sort_is_stable_brute_force():
    rng = np.random.default_rng(1)
    src = rng.integers(0, 17, 500).astype(np.uint32)
    dst = rng.integers(0, 17, 500).astype(np.uint32)
    w = np.arange(500, dtype=np.float32)
    s, d, ww = O.count_sort(src, dst, w, 17, mode=2, move_weights=True)
    order = sorted(range(500), key=lambda i: int(src[i]))  # Python's sort is stable
    np.testing.assert_array_equal(ww, w[order])
    s3, d3, w3 = O.count_sort(src, dst, w, 17, mode=3, move_weights=True)
    order3 = sorted(range(500), key=lambda i: (int(src[i]), int(dst[i])))
    np.testing.assert_array_equal(w3, w[order3])


def test_dedupe_set_semantics():
    rng = np.random.default_rng(2)
    src = rng.integers(0, 20, 800).astype(np.uint32)
    dst = rng.integers(0, 20, 800).astype(np.uint32)
    fwd, _ = O.build_graph(src, dst, None, 20, dedupe=True)
    pairs = sorted({(int(a), int(b)) for a, b in zip(src, dst) if a != b})
    assert fwd.E == len(pairs)
    np.testing.assert_array_equal(fwd.src, [p[0] for p in pairs])
    np.testing.assert_array_equal(fwd.dst, [p[1] for p in pairs])


# ---------------------------------------------------------------- PageRank / SpMV

@pytest.mark.parametrize("name", GRAPHS)
def test_oracle_pr_f32_bit_exact_with_reference(golden, name):
    z = golden(name)
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    r1 = O.pr_pull_f32(fwd, inv, 1, 0.0, nthreads=2)
    r20 = O.pr_pull_f32(fwd, inv, 20, 0.0, nthreads=2)
    np.testing.assert_array_equal(r1["ranks"], z["pr_r1"])
    np.testing.assert_array_equal(r20["ranks"], z["pr_r20"])
    assert r20["error_total"] == pytest.approx(float(z["pr_err20"]), rel=1e-9)
    re = O.pr_pull_f32(fwd, inv, 100, 1e-4, nthreads=2)
    assert re["iterations"] == int(z["pr_eps_iters"])
    np.testing.assert_array_equal(re["ranks"], z["pr_eps_ranks"])


@pytest.mark.parametrize("name", GRAPHS)
def test_oracle_spmv_bit_exact_with_reference(golden, name):
    z = golden(name)
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    x, y = O.spmv_pull_f32(fwd, inv, 3, nthreads=2)
    np.testing.assert_array_equal(x, z["spmv_x"])
    np.testing.assert_array_equal(y, z["spmv_y3"])


@pytest.mark.parametrize("s", RMAT_SMALL)
def test_oracle_weighted_spmv_bit_exact_with_reference(golden, s):
    z = golden(f"rmat_s{s}")
    ws, wd, ww = z["wedges_src"], z["wedges_dst"], z["wedges_w"]
    fwd, inv = O.build_graph(ws, wd, ww, int(z["m1_N"]), mode=1, move_weights=True)
    np.testing.assert_array_equal(inv.w, z["wm1_inv_w"])
    x, y = O.spmv_pull_f32(fwd, inv, 1, weights=inv.w, nthreads=2)
    np.testing.assert_array_equal(y, z["wspmv_y1"])


@pytest.mark.parametrize("name", GRAPHS)
def test_fp64_ref_oracle_close_to_reference_f32(golden, name):
    z = golden(name)
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    r64, it, _ = O.pr_f64(fwd.N, fwd.out_degree, O.row_ptr_from(inv), inv.dst, mode=0, iterations=20)
    assert it == 20
    assert l1(r64, z["pr_r20"]) < 1e-6


def test_fp64_ref_matches_dense_fixed_point(golden):
    z = golden("fixture_v500_e500")
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    rp = O.row_ptr_from(inv)
    r64, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=0, iterations=400)
    dense = O.dense_ref_fixed_point(fwd.N, fwd.out_degree, rp, inv.dst)
    assert np.abs(r64 - dense).max() < 1e-14


def test_fp64_stochastic_matches_dense_solve_and_sums_to_one(golden):
    z = golden("rmat_s8")
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    rp = O.row_ptr_from(inv)
    r, it, err = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=500, epsilon=1e-15)
    dense = O.dense_stochastic(fwd.N, fwd.out_degree, rp, inv.dst)
    assert abs(r.sum() - 1.0) < 1e-12
    assert np.abs(r - dense).max() < 1e-13
    # personalised: linear in the teleport vector
    rng = np.random.default_rng(3)
    t1 = rng.random(fwd.N); t1 /= t1.sum()
    t2 = rng.random(fwd.N); t2 /= t2.sum()
    p1, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=500, epsilon=1e-15, teleport=t1)
    p2, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=500, epsilon=1e-15, teleport=t2)
    p12, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=1, iterations=500, epsilon=1e-15,
                         teleport=0.3 * t1 + 0.7 * t2)
    assert np.abs(p12 - (0.3 * p1 + 0.7 * p2)).max() < 1e-13


def _cycle(n):
    src = np.arange(n, dtype=np.uint32)
    dst = ((np.arange(n) + 1) % n).astype(np.uint32)
    return O.build_graph(src, dst, None, n, mode=1)


@pytest.mark.parametrize("k", [1, 5, 20])
def test_cycle_closed_forms(k):
    n = 37
    fwd, inv = _cycle(n)
    rp = O.row_ptr_from(inv)
    r, _, _ = O.pr_f64(n, fwd.out_degree, rp, inv.dst, mode=0, iterations=k)
    # REF: r_k = 1 - 0.85^(k+1) at every vertex (r_0 = 0.15), output / N
    np.testing.assert_allclose(r, (1 - 0.85 ** (k + 1)) / n, rtol=1e-13)
    f32 = O.pr_pull_f32(fwd, inv, k, 0.0)
    np.testing.assert_allclose(f32["ranks"], (1 - 0.85 ** (k + 1)) / n, rtol=1e-5)
    s, _, _ = O.pr_f64(n, fwd.out_degree, rp, inv.dst, mode=1, iterations=k)
    np.testing.assert_allclose(s, 1.0 / n, rtol=1e-13)


def test_planted_mistakes_are_caught(golden):
    """Mistakes planted in the oracle's inputs fail the closed forms above."""
    n = 37
    fwd, inv = _cycle(n)
    rp = O.row_ptr_from(inv)
    good = (1 - 0.85 ** 6) / n
    # wrong damping
    r, _, _ = O.pr_f64(n, fwd.out_degree, rp, inv.dst, mode=0, iterations=5, damp=0.8)
    assert not np.allclose(r, good, rtol=1e-6)
    # transposed operand on a non-symmetric graph: forward rows instead of in-rows
    z = golden("fixture_v1024_e4096")
    f, i = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    wrong, _, _ = O.pr_f64(f.N, f.out_degree, O.row_ptr_from(f), f.dst, mode=0, iterations=20)
    assert l1(wrong, z["pr_r20"]) > 1e-3
    # dropped dangling term breaks the sum-to-one invariant
    z8 = golden("rmat_s8")
    f8, i8 = O.build_graph(z8["edges_src"], z8["edges_dst"], None, int(z8["m1_N"]), mode=1)
    assert (f8.out_degree == 0).any()
    r_ref, _, _ = O.pr_f64(f8.N, f8.out_degree, O.row_ptr_from(i8), i8.dst, mode=0, iterations=50)
    assert abs(r_ref.sum() - 1.0) > 1e-3


def test_adjoint_matches_finite_differences_and_dense():
    rng = np.random.default_rng(7)
    n = 60
    src = rng.integers(0, n, 400).astype(np.uint32)
    dst = rng.integers(0, n, 400).astype(np.uint32)
    fwd, inv = O.build_graph(src, dst, None, n, dedupe=True)
    rp = O.row_ptr_from(inv)
    wraw = np.exp(rng.standard_normal(inv.E))
    g = rng.standard_normal(n)
    t = rng.random(n); t /= t.sum()
    p, lam, gw, gt = O.adjoint_f64(n, rp, inv.dst, wraw, g, teleport=t)

    def loss(wr, tt):
        S = np.zeros(n)
        np.add.at(S, inv.dst, wr)
        P = wr / S[inv.dst]
        pp = O.dense_stochastic(n, fwd.out_degree, rp, inv.dst, teleport=tt, weights=P)
        return float(g @ pp)

    assert loss(wraw, t) == pytest.approx(float(g @ p), rel=1e-10)
    h = 1e-6
    for e in rng.choice(inv.E, 12, replace=False):
        wp, wm = wraw.copy(), wraw.copy()
        wp[e] += h
        wm[e] -= h
        fd = (loss(wp, t) - loss(wm, t)) / (2 * h)
        assert gw[e] == pytest.approx(fd, rel=1e-5, abs=1e-10)
    for v in rng.choice(n, 6, replace=False):
        tp, tm = t.copy(), t.copy()
        tp[v] += h
        tm[v] -= h
        fd = (loss(wraw, tp) - loss(wraw, tm)) / (2 * h)
        assert gt[v] == pytest.approx(fd, rel=1e-6, abs=1e-10)


def test_loss_diff_matches_dense_solves():
    rng = np.random.default_rng(8)
    n = 80
    src = rng.integers(0, n, 600).astype(np.uint32)
    dst = rng.integers(0, n, 600).astype(np.uint32)
    fwd, inv = O.build_graph(src, dst, None, n, dedupe=True)
    rp = O.row_ptr_from(inv)
    wraw = np.exp(rng.standard_normal(inv.E))
    g = rng.standard_normal(n)

    def dense(wr):
        S = np.zeros(n)
        np.add.at(S, inv.dst, wr)
        return float(g @ O.dense_stochastic(n, fwd.out_degree, rp, inv.dst, weights=wr / S[inv.dst]))

    wp, wm = wraw * 1.01, wraw.copy()
    wm[3] *= 0.5
    got = O.loss_diff(n, rp, inv.dst, wp, wm, g, iterations=400)
    assert got == pytest.approx(dense(wp) - dense(wm), rel=1e-9, abs=1e-15)
    L, p = O.loss_f64(n, rp, inv.dst, wraw, g, iterations=400)
    assert L == pytest.approx(dense(wraw), rel=1e-12)
    assert p.sum() == pytest.approx(1.0, abs=1e-12)


def test_adjoint_fd_fixture_pins_fp64_adjoint(golden):
    """R-MAT scale 20 differentiable workload: the oracle's FP64 adjoint
    against the stored FP64 central differences (forward solves only)."""
    fx = golden("adjoint_fd_s20")
    s, d, w, N, g = O.weighted_rmat_workload(int(fx["scale"]), int(fx["edge_factor"]), int(fx["seed"]))
    assert (N, s.size) == (int(fx["N"]), int(fx["E"]))
    fwd, inv = O.build_graph(s, d, w, N, mode=1, move_weights=True)
    rp = O.row_ptr_from(inv)
    p, lam, gw, gt = O.adjoint_f64(N, rp, inv.dst, inv.w.astype(np.float64), g, iterations=2000, tol=1e-15)
    assert float(g @ p) == pytest.approx(float(fx["loss"]), rel=1e-12)
    key = inv.dst.astype(np.uint64) * np.uint64(N) + inv.src.astype(np.uint64)
    order = np.argsort(key)
    fk = fx["u"].astype(np.uint64) * np.uint64(N) + fx["v"].astype(np.uint64)
    at = order[np.searchsorted(key[order], fk)]
    assert np.array_equal(key[at], fk)
    rel = np.abs(gw[at] - fx["fd"]) / np.abs(fx["fd"])
    assert rel.max() <= 2e-5
    # whole-gradient directional derivative (edge-list order xi)
    xi_list = np.random.default_rng(int(fx["seed"]) + 400 + int(fx["scale"])).standard_normal(s.size)
    kl = s.astype(np.uint64) * np.uint64(N) + d.astype(np.uint64)
    ol = np.argsort(kl)
    xi = xi_list[ol][np.searchsorted(kl[ol], key)]
    dd = float(np.sum(gw * xi * inv.w.astype(np.float64)))
    assert dd == pytest.approx(float(fx["dir_fd"]), rel=1e-5)


VARIANT_GRAPHS = ["fixture_v500_e500", "fixture_v1024_e4096", "fixture_v8192_e32768", "fixture_v1000_e10000",
                  "fixture_v529_e3500", "rmat_s8", "rmat_s10", "rmat_s12"]


@pytest.mark.parametrize("name", VARIANT_GRAPHS)
def test_fixed_point_oracle_matches_reference_variants(golden, name):
    """pageRankPullFixedPoint{64,32,16,8}BitGraphCSR (pushpull 2, 9, 10, 11):
    bit-exact ranks and identical stopping iterations."""
    z, v = golden(name), golden(f"variants_{name}")
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    rp = O.row_ptr_from(inv)
    for bits, code in ((64, 2), (32, 9), (16, 10), (8, 11)):
        r, _, err = O.pr_pull_fixed(fwd.N, fwd.out_degree, rp, inv.dst, bits, 20, 0.0)
        np.testing.assert_array_equal(r, v[f"pp{code}_r20"])
        assert err == pytest.approx(float(v[f"pp{code}_err20"]), rel=1e-9)
        r, it, _ = O.pr_pull_fixed(fwd.N, fwd.out_degree, rp, inv.dst, bits, 100, 1e-4)
        assert it == int(v[f"pp{code}_eps_iters"])
        np.testing.assert_array_equal(r, v[f"pp{code}_eps_ranks"])


@pytest.mark.parametrize("name", VARIANT_GRAPHS)
def test_reference_push_variants_equal_pull_sums(golden, name):
    """Push (1) sums the same contributions as pull; push fixed point (3)
    the same integers as pull fixed point 64 (2): the reference's own
    outputs agree (one thread, so float atomics run in edge order)."""
    z, v = golden(name), golden(f"variants_{name}")
    assert O_l1(v["pp1_r20"], z["pr_r20"]) <= 1e-6
    np.testing.assert_array_equal(v["pp3_r20"], v["pp2_r20"])
    np.testing.assert_array_equal(v["pp3_eps_ranks"], v["pp2_eps_ranks"])


RING_GRAPHS = ["ring_rmat_s10", "ring_64"]  # no dangling vertex: quantisation min > 0, and min == max


def _variant_csr(golden, name):
    """In-CSR of a golden graph; the ring graphs carry their own edges."""
    v = golden(f"variants_{name}")
    if name in RING_GRAPHS:
        s, d = v["edges_src"], v["edges_dst"]
        N = int(max(s.max(), d.max())) + 1
    else:
        z = golden(name)
        s, d, N = z["edges_src"], z["edges_dst"], int(z["m1_N"])
    fwd, inv = O.build_graph(s, d, None, N, mode=1)
    return v, fwd, inv, O.row_ptr_from(inv)


@pytest.mark.parametrize("name", VARIANT_GRAPHS + RING_GRAPHS)
def test_quantized_oracle_matches_reference_variants(golden, name):
    """pageRankPullQuant{32,16,8}BitGraphCSR (pushpull 4, 12, 13): bit-exact
    ranks, error and stopping iterations, including the graphs without
    dangling vertices where the 32-bit quantisation wraps past 2^32 and the
    16/8-bit one saturates (ring_rmat_s10), and where every contribution is
    equal (ring_64: scale 1)."""
    v, fwd, inv, rp = _variant_csr(golden, name)
    for bits, code in ((32, 4), (16, 12), (8, 13)):
        r, _, err = O.pr_pull_quant(fwd.N, fwd.out_degree, rp, inv.dst, bits, 20, 0.0)
        np.testing.assert_array_equal(r, v[f"pp{code}_r20"])
        assert err == pytest.approx(float(v[f"pp{code}_err20"]), rel=1e-9, abs=1e-300)
        r, it, _ = O.pr_pull_quant(fwd.N, fwd.out_degree, rp, inv.dst, bits, 100, 1e-4)
        assert it == int(v[f"pp{code}_eps_iters"])
        np.testing.assert_array_equal(r, v[f"pp{code}_eps_ranks"])


@pytest.mark.parametrize("name", RING_GRAPHS)
def test_fixed_point_oracle_on_graphs_without_dangling_vertices(golden, name):
    v, fwd, inv, rp = _variant_csr(golden, name)
    for bits, code in ((64, 2), (32, 9), (16, 10), (8, 11)):
        r, _, _ = O.pr_pull_fixed(fwd.N, fwd.out_degree, rp, inv.dst, bits, 20, 0.0)
        np.testing.assert_array_equal(r, v[f"pp{code}_r20"])


def test_ring_golden_reaches_the_wrap_and_saturation():
    """The first iteration of ring_rmat_s10 quantises values past the
    width's range: c / scale + 0.5 >= 2^32 for the 32-bit variant (the u32
    wrap) and above 65535 / 255 for 16 / 8 bits (CLAMP), so the bit-exact
    match above pins both rules."""
    with np.load(os.path.join(GOLDEN, "variants_ring_rmat_s10.npz"), allow_pickle=False) as z:
        s, d = z["edges_src"], z["edges_dst"]
    N = int(max(s.max(), d.max())) + 1
    fwd, _ = O.build_graph(s, d, None, N, mode=1)
    assert fwd.out_degree.min() > 0
    c = (np.float32(1) - np.float32(0.85)) / fwd.out_degree.astype(np.float32)
    spread = np.float64(c.max()) - np.float64(c.min())
    for rng, as_float in ((4294967295.0, False), (65535.0, True), (255.0, True)):
        scale = spread / rng
        if as_float:
            scale = np.float64(np.float32(scale))
        x = c.astype(np.float64) / scale + 0.5
        assert (x >= rng + 1).sum() > 100
    assert (c.astype(np.float64) / (spread / 4294967295.0) + 0.5 >= 2.0**32).sum() > 100


@pytest.mark.parametrize("name", VARIANT_GRAPHS + RING_GRAPHS)
def test_data_driven_pull_oracle_matches_reference(golden, name):
    """pageRankDataDrivenPullGraphCSR (pushpull 6): bit-exact ranks, error
    and stopping iteration (sequential float row sums, as the reference)."""
    v, fwd, inv, rp = _variant_csr(golden, name)
    frp = O.row_ptr_from(fwd)
    r, _, err = O.pr_pull_dd(fwd.N, fwd.out_degree, rp, inv.dst, frp, fwd.dst, 20, 0.0)
    np.testing.assert_array_equal(r, v["pp6_r20"])
    assert err == float(v["pp6_err20"])
    r, it, err = O.pr_pull_dd(fwd.N, fwd.out_degree, rp, inv.dst, frp, fwd.dst, 100, 1e-4)
    assert it == int(v["pp6_eps_iters"]) and err == float(v["pp6_eps_err"])
    np.testing.assert_array_equal(r, v["pp6_eps_ranks"])


def test_data_driven_pull_is_ref_at_epsilon_zero_and_differs_above(golden):
    """With epsilon 0 every worklist row is updated and a row leaves the
    worklist only when its in-neighbours stop changing, so the data-driven
    pull is the plain pull; with epsilon > 0 rows below the threshold keep
    their rank and the outputs differ."""
    v, fwd, inv, rp = _variant_csr(golden, "rmat_s12")
    np.testing.assert_array_equal(v["pp6_r20"], golden("rmat_s12")["pr_r20"])
    assert not np.array_equal(v["pp6_eps_ranks"], golden("rmat_s12")["pr_eps_ranks"])


@pytest.mark.parametrize("name", VARIANT_GRAPHS + RING_GRAPHS)
def test_data_driven_push_oracles_match_reference(golden, name):
    """pageRankDataDriven{Push,PullPush}GraphCSR (pushpull 7, 8) as one
    reference thread runs them: bit-exact ranks, error and iterations."""
    v, fwd, inv, rp = _variant_csr(golden, name)
    frp = O.row_ptr_from(fwd)
    for code, pp in ((7, False), (8, True)):
        r, _, err = O.pr_dd_push(fwd.N, fwd.out_degree, rp, inv.dst, frp, fwd.dst, pp, 20, 0.0)
        np.testing.assert_array_equal(r, v[f"pp{code}_r20"])
        assert err == float(v[f"pp{code}_err20"])
        r, it, err = O.pr_dd_push(fwd.N, fwd.out_degree, rp, inv.dst, frp, fwd.dst, pp, 100, 1e-4)
        assert it == int(v[f"pp{code}_eps_iters"]) and err == float(v[f"pp{code}_eps_err"])
        np.testing.assert_array_equal(r, v[f"pp{code}_eps_ranks"])


@pytest.mark.parametrize("name", ["fixture_v1024_e4096", "fixture_v529_e3500", "rmat_s12"])
def test_data_driven_push_limits(golden, name):
    """The fixed points the reference's own algorithms approach as epsilon
    shrinks: the push form (7) the self-loop-free limit of dd_push_limit (the
    fixtures hold self loops, whose later pushes the reference drops; the
    ordinary fixed point is 1e-3 away there), the pull-push form (8) the
    ordinary REF fixed point.  The GPU forms are held to the same limits."""
    v, fwd, inv, rp = _variant_csr(golden, name)
    frp = O.row_ptr_from(fwd)
    lim = O.dd_push_limit(fwd.N, fwd.out_degree, rp, inv.dst)
    fix, _, _ = O.pr_f64(fwd.N, fwd.out_degree, rp, inv.dst, mode=0, iterations=2000, epsilon=1e-15)
    r7, _, _ = O.pr_dd_push(fwd.N, fwd.out_degree, rp, inv.dst, frp, fwd.dst, False, 3000, 1e-7)
    r8, _, _ = O.pr_dd_push(fwd.N, fwd.out_degree, rp, inv.dst, frp, fwd.dst, True, 3000, 1e-7)
    assert O_l1(r7, lim) <= 1e-6
    assert O_l1(r8, fix) <= 1e-6
    if name.startswith("fixture"):
        assert O_l1(lim, fix) > 1e-4  # the self-loop rule matters on these graphs


def test_quantized_goldens_tell_widths_apart(golden):
    v, fwd, inv, rp = _variant_csr(golden, "rmat_s10")
    r16, _, _ = O.pr_pull_quant(fwd.N, fwd.out_degree, rp, inv.dst, 16, 20, 0.0)
    assert not np.array_equal(r16, v["pp13_r20"]) and not np.array_equal(r16, v["pp4_r20"])
    r8, _, _ = O.pr_pull_quant(fwd.N, fwd.out_degree, rp, inv.dst, 8, 19, 0.0)
    assert not np.array_equal(r8, v["pp13_r20"])


def O_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).sum())


def test_fixed_point_goldens_tell_widths_and_iterations_apart(golden):
    # a width mixed up, or an iteration too few, would not pass the bit-exact checks
    z, v = golden("rmat_s10"), golden("variants_rmat_s10")
    fwd, inv = O.build_graph(z["edges_src"], z["edges_dst"], None, int(z["m1_N"]), mode=1)
    rp = O.row_ptr_from(inv)
    r16, _, _ = O.pr_pull_fixed(fwd.N, fwd.out_degree, rp, inv.dst, 16, 20, 0.0)
    assert not np.array_equal(r16, v["pp9_r20"])  # a 16-bit run is not the 32-bit variant
    r32, _, _ = O.pr_pull_fixed(fwd.N, fwd.out_degree, rp, inv.dst, 32, 19, 0.0)
    assert not np.array_equal(r32, v["pp9_r20"])  # nor one iteration fewer


# ---------------------------------------------------------------- SpMV variants

def _ref_spmv_graph(tmp_path, z, weighted):
    R = O.ref(1, int(weighted))
    path = str(tmp_path / ("g.wbin" if weighted else "g.bin"))
    if weighted:
        O.write_bin(path, z["wedges_src"], z["wedges_dst"], z["wedges_w"])
    else:
        O.write_bin(path, z["edges_src"], z["edges_dst"])
    g = R.refd_build(path.encode(), 1, 0, 0, int(weighted), 1)
    N, E = R.refd_num_vertices(g), R.refd_num_edges(g)
    a = [np.zeros(N, np.uint32) for _ in range(3)] + [np.zeros(E, np.uint32) for _ in range(2)] + \
        [np.zeros(E, np.float32)]
    R.refd_export(g, 1, *[x.ctypes.data for x in a])
    fod = np.zeros(N, np.uint32)
    R.refd_export(g, 0, fod.ctypes.data, None, None, None, None, None)
    rp = np.zeros(N + 1, np.uint64)
    np.cumsum(a[0], out=rp[1:])
    return R, g, N, fod, rp, a[4], a[5]


@pytest.mark.parametrize("name", ["rmat_s10", "rmat_s12"])
@pytest.mark.parametrize("weighted", [False, True])
def test_spmv_fixed_oracle_matches_reference(golden, tmp_path, name, weighted):
    """SPMVPullFixedPointGraphCSR (2) and SPMVPushFixedPointGraphCSR (3):
    bit-exact inputs and outputs after 1 and 3 accumulating passes."""
    R, g, N, fod, rp, col, w = _ref_spmv_graph(tmp_path, golden(name), weighted)
    for code, bits in ((2, 32), (3, 64)):
        for it in (1, 3):
            y, x = np.zeros(N, np.float32), np.zeros(N, np.float32)
            R.refd_spmv_variant(g, code, it, 1, y, x)
            xo, yo = O.spmv_fixed(N, fod, rp, col, w if weighted else None, bits, it)
            np.testing.assert_array_equal(xo, x)
            np.testing.assert_array_equal(yo, y, err_msg=f"code {code} iterations {it}")
    # the push forms sum the pull's terms: one reference thread gives identical floats
    y0, y1, xd = np.zeros(N, np.float32), np.zeros(N, np.float32), np.zeros(N, np.float32)
    R.refd_spmv_variant(g, 0, 3, 1, y0, xd)
    R.refd_spmv_variant(g, 1, 3, 1, y1, xd)
    np.testing.assert_array_equal(y0, y1)
    R.refd_free(g)


def test_spmv_fixed_64_product_uses_the_high_bits():
    """((u128) w * x) >> 32 needs the high half: weights and inputs near 2^32
    (a weight of ~1 times an input of ~1) tell a 64-bit-only product apart."""
    N = 3
    rp = np.array([0, 1, 2, 3], np.uint64)
    col = np.array([1, 2, 0], np.uint32)
    w = np.array([0.9, 1.0, 0.5], np.float32)
    _, y = O.spmv_fixed(N, np.array([1, 1, 1], np.uint32), rp, col, w, 64, 1)
    np.testing.assert_allclose(y, [0.9, 1.0, 0.5], rtol=1e-6)


# ---------------------------------------------------------------- relabel

LMODES = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14)


def _ref_relabel(path, dflag, levels, fnamel=None, sort=1):
    """Labels and forward CSR arrays of the reference's pipeline with relabel levels."""
    R = O.ref(1, 0)
    g = R.refd_build_relabel(path.encode(), sort, dflag, 0, 0, 1, *levels, fnamel.encode() if fnamel else None)
    N, E = R.refd_num_vertices(g), R.refd_num_edges(g)
    lab = np.zeros(N, np.uint32)
    R.refd_labels(g, lab.ctypes.data, None)
    src, dst = np.zeros(E, np.uint32), np.zeros(E, np.uint32)
    R.refd_export(g, 0, None, None, None, src.ctypes.data, dst.ctypes.data, None)
    R.refd_free(g)
    return lab, src, dst


@pytest.mark.parametrize("name", ["rmat_s10", "fixture_v1024_e4096"])
@pytest.mark.parametrize("dflag", [0, 1])
def test_relabel_restatement_matches_reference(golden, tmp_path, name, dflag):
    """reorderGraphProcess's orders (reorder.c:296-1164, 1483-1616) for every
    lmode, and the pipeline around them: the renamed edges sorted again
    stably (graphCSR.c:268-290)."""
    z = golden(name)
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    lab0, src, dst = _ref_relabel(path, dflag, (0, 0, 0))
    N = lab0.size
    avg = (src.size if dflag else len(z["edges_src"])) // N
    for lm in LMODES:
        lab, rs, rd = _ref_relabel(path, dflag, (lm, 0, 0))
        mine = O.relabel_labels(src, dst, N, avg, lm)
        np.testing.assert_array_equal(mine, lab, err_msg=f"lmode {lm}")
        ns, nd = mine[src].astype(np.uint64), mine[dst].astype(np.uint64)
        order = np.argsort(ns * np.uint64(N) + nd, kind="stable")
        np.testing.assert_array_equal(ns[order], rs, err_msg=f"lmode {lm}")
        np.testing.assert_array_equal(nd[order], rd, err_msg=f"lmode {lm}")


def test_relabel_levels_compose_and_file_labels(golden, tmp_path):
    z = golden("rmat_s10")
    path = str(tmp_path / "g.bin")
    O.write_bin(path, z["edges_src"], z["edges_dst"])
    lab0, src, dst = _ref_relabel(path, 0, (0, 0, 0))
    N = lab0.size
    avg = len(z["edges_src"]) // N
    total = np.arange(N, dtype=np.uint32)
    s, d = src, dst
    for lm in (4, 1, 8):
        step = O.relabel_labels(s, d, N, avg, lm)
        s, d = step[s], step[d]
        total = step[total]
    lab, _, _ = _ref_relabel(path, 0, (4, 1, 8))
    np.testing.assert_array_equal(total, lab)
    perm = np.random.default_rng(5).permutation(N).astype(np.uint32)
    fl = str(tmp_path / "labels.txt")
    np.savetxt(fl, perm, fmt="%u")
    lab, rs, _ = _ref_relabel(path, 0, (11, 0, 0), fnamel=fl)
    np.testing.assert_array_equal(lab, perm)
    np.testing.assert_array_equal(np.sort(perm[src]), rs)


def test_mt19937_stream():
    """The MT19937 stream of the random-degree order: the reference's random
    relabel (pinned above) starts from it; the generator is the published one
    (equidistribution: outputs cover the 32 bits evenly)."""
    x = np.array(O.mt19937_uint32(27491095, 20000), np.uint64)
    assert len(set(x.tolist())) == x.size
    bits = np.array([(x >> np.uint64(b)) & np.uint64(1) for b in range(32)]).mean(axis=1)
    assert np.all(np.abs(bits - 0.5) < 0.02)


# ---------------------------------------------------------------- generator

def test_rmat_permutation_is_bijection():
    for scale in (1, 2, 5, 11, 14):
        img = {O.rmat_permute(x, scale, 99) for x in range(1 << scale)}
        assert img == set(range(1 << scale))


def test_rmat_quadrant_frequencies():
    s, d = O.rmat(1, 1 << 15, 5)  # scale 1: one level, ids permuted on {0,1}
    inv = {O.rmat_permute(x, 1, 5): x for x in (0, 1)}
    u = np.vectorize(inv.get)(s)
    v = np.vectorize(inv.get)(d)
    M = s.size
    freq = np.array([((u == a) & (v == b)).sum() for a in (0, 1) for b in (0, 1)]) / M
    np.testing.assert_allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.01)


def test_rmat_dedupe_keeps_first_in_order():
    s, d = O.rmat(6, 16, 11)
    ks, kd, _ = O.dedupe_keep_first(s, d)
    seen, exp = set(), []
    for a, b in zip(s.tolist(), d.tolist()):
        if a != b and (a, b) not in seen:
            seen.add((a, b))
            exp.append((a, b))
    assert list(zip(ks.tolist(), kd.tolist())) == exp


def test_rmat_s16_regenerates_golden(golden):
    import hashlib
    z = golden("rmat_s16")
    s, d, N = O.rmat_graph(16, 16, 20261203 + 16)
    assert N == int(z["m1_N"]) and s.size == int(z["m1_E"])
    assert hashlib.sha256(np.ascontiguousarray(np.stack([s, d])).tobytes()).hexdigest() == str(z["edges_sha"])


# ---------------------------------------------------------------- native libraries

def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gbx_[a-z0-9_]+)\s*\(", text)))


def test_hip_library_exports_every_declared_symbol():
    import ctypes
    import graphblox_amd as gbx

    if not os.path.exists(gbx.LIB_HIP):
        pytest.fail("libgbx_hip.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(gbx.LIB_HIP)
    names = _declared("gbx.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(gbx.exported_symbols())
    assert b"gfx950" in gbx.lib().gbx_version()
