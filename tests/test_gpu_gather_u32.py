"""The FPGA overlay's Single.PR integer pull on the GPU
(gbx_gather_accumulate_u32): for every row v, aux2[N + v] = the sum of
aux1[N + u] over v's edges with ALU_ACC's unsigned 32-bit add
(engine_alu_ops_kernel.sv:104-112), on 2N-word buffers as the host maps them
(buffer_mapping.PR.cpp:2-68, test_GraphBlox_USER_MANAGED.cpp:775-804).

* aux1 = 1 on the second half gives every row's degree: the reference's own
  host check (test_GraphBlox_USER_MANAGED.cpp:518-528), against the
  reference's transpose degrees (golden) for the DIRECTED rows;
* random u32 values, including ones whose sums wrap past 2^32, against a
  brute-force numpy uint32 sum over the same CSR, both directions;
* the first halves: aux2[0, N) is left as it was."""
import numpy as np
import pytest

import graphblox_amd as gbx

pytestmark = pytest.mark.gpu


def _brute(N, rows, vals):
    """sum over each row's columns of vals[col], modulo 2^32"""
    od, idx, col = rows["out_degree"], rows["edges_idx"], rows["dst"]
    out = np.zeros(N, np.uint32)
    for v in np.flatnonzero(od):
        b = int(idx[v])
        out[v] = np.add.reduce(vals[col[b:b + int(od[v])] & 0x3FFFFFFF], dtype=np.uint32)
    return out


@pytest.mark.parametrize("name", ["fixture_v1024_e4096", "rmat_s12"])
def test_all_ones_gives_the_degree(gpu_ctx, golden, name):
    z = golden(name)
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"], sort_mode=1)
    N = g.N
    aux1 = np.zeros(2 * N, np.uint32)
    aux1[N:] = 1  # initialize_PR_auxiliary_struct (test_GraphBlox_USER_MANAGED.cpp:775-804)
    aux2 = g.gather_accumulate_u32(aux1)
    np.testing.assert_array_equal(aux2[N:], z["m1_inv_out_degree"])  # in-degree of every vertex
    assert not aux2[:N].any()
    aux2 = g.gather_accumulate_u32(aux1, direction=gbx.PULL_FWD)
    np.testing.assert_array_equal(aux2[N:], z["m1_fwd_out_degree"])
    g.free()


@pytest.mark.parametrize("scale", [10, 16])
def test_random_u32_values_wrap_like_the_alu(gpu_ctx, scale):
    g = gpu_ctx.rmat_graph(scale, 16, 5)
    N = g.N
    rng = np.random.default_rng(scale)
    aux1 = rng.integers(0, 2**32, 2 * N, dtype=np.uint64).astype(np.uint32)
    aux1[N + rng.integers(0, N, 64)] = np.uint32(0xFFFFFFFF)
    before = rng.integers(0, 2**32, 2 * N, dtype=np.uint64).astype(np.uint32)
    for direction, side in ((gbx.PULL_IN, True), (gbx.PULL_FWD, False)):
        rows = g.export(inverse=side)
        want = _brute(N, rows, aux1[N:])
        # sums do wrap: the exact sums exceed 2^32 for most rows with edges
        exact = np.zeros(N, np.uint64)
        np.add.at(exact, rows["src"] & 0x3FFFFFFF, aux1[N:][rows["dst"] & 0x3FFFFFFF].astype(np.uint64))
        assert (exact >= 2**32).sum() > N // 4
        got = g.gather_accumulate_u32(aux1, before, direction=direction)
        np.testing.assert_array_equal(got[N:], want)
        np.testing.assert_array_equal(got[N:], (exact % 2**32).astype(np.uint32))
        np.testing.assert_array_equal(got[:N], before[:N])
    g.free()


def test_buffers_must_hold_2n_words(gpu_ctx, golden):
    z = golden("fixture_v500_e500")
    g = gpu_ctx.graph(z["edges_src"], z["edges_dst"])
    with pytest.raises(ValueError):
        g.gather_accumulate_u32(np.ones(g.N, np.uint32))
    g.free()
