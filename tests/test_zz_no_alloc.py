"""gbx_spmv_pull_noalloc, the SpMV entry point that allocates nothing: through
the C ABI on device vectors, against gbx_spmv_pull on the same graph and
input (bit-identical: the same plan, kernels and order of additions), two
calls in a row with the second on a smaller graph (the scratch array keeps
the first call's contents), weighted and unweighted, both directions, and
its refusals (no plan yet, NULL vectors)."""
import ctypes as C

import numpy as np
import pytest

import graphblox_amd as gbx

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libgbx_hip.so runs on: opened by its soname once the
    library is loaded, the dynamic linker hands back the object it bound
    libgbx_hip.so to."""
    gbx.lib()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    return H


class DeviceArray:
    """float32 device memory of the test (the entry point under test allocates none)."""

    def __init__(self, a):
        self.H = _hip()
        a = np.ascontiguousarray(a, np.float32)
        self.n = a.size
        self.p = C.c_void_p()
        assert self.H.hipMalloc(C.byref(self.p), max(4, a.nbytes)) == 0
        assert self.H.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty(self.n, np.float32)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    @property
    def ptr(self):
        return self.p.value

    def free(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = C.c_void_p()


def _run_noalloc(ctx, g, x, iterations, weighted, direction):
    dx = DeviceArray(x)
    dy = DeviceArray(np.full(g.N, 7.0, np.float32))  # overwritten: y starts at 0
    g.spmv_noalloc(dx.ptr, dy.ptr, iterations=iterations, weighted=weighted, direction=direction)
    ctx.sync()
    y = dy.get()
    dx.free()
    dy.free()
    return y


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("direction", [gbx.PULL_IN, gbx.PULL_FWD])
def test_noalloc_spmv_matches_gbx_spmv_pull(gpu_ctx, weighted, direction):
    # hub rows cut into chunks of 1024 edges, combined through the scratch
    # array; a graph with many such rows, then a smaller one
    saved = gpu_ctx.tuning()
    t = gpu_ctx.tuning()
    t.hub_chunk = 1024
    t.bin_thresholds[0] = 1024
    gpu_ctx.set_tuning(t)
    try:
        _check_pair(gpu_ctx, weighted, direction)
    finally:
        gpu_ctx.set_tuning(saved)


def _check_pair(gpu_ctx, weighted, direction):
    big = gpu_ctx.rmat_graph(18, 16, 9, weight_kind=1 if weighted else 0, weight_seed=3,
                             flags=gbx.GRAPH_WEIGHTED if weighted else 0)
    small = gpu_ctx.rmat_graph(11, 8, 4, weight_kind=1 if weighted else 0, weight_seed=5,
                               flags=gbx.GRAPH_WEIGHTED if weighted else 0)
    rng = np.random.default_rng(7)
    s = big.session(direction=direction)
    hub = s.plan_info()[0]  # class 0: hub rows, cut into chunks
    s.free()
    assert hub["lanes"] == 256 and hub["edges"] > hub["rows"] * 1024, hub  # rows of several chunks
    for g, iters in ((big, 3), (small, 2)):
        print(f"[gpu] noalloc SpMV: N={g.N} E={g.E} weighted={weighted} direction={direction}", flush=True)
        x = rng.random(g.N).astype(np.float32)
        _, y_ref, _ = g.spmv(iterations=iters, weighted=weighted, direction=direction, x=x)  # builds the plan
        y = _run_noalloc(gpu_ctx, g, x, iters, weighted, direction)
        np.testing.assert_array_equal(y, y_ref)
    # the first graph again after the smaller one: nothing left over in the scratch array
    x = rng.random(big.N).astype(np.float32)
    _, y_ref, _ = big.spmv(iterations=1, weighted=weighted, direction=direction, x=x)
    np.testing.assert_array_equal(_run_noalloc(gpu_ctx, big, x, 1, weighted, direction), y_ref)
    big.free()
    small.free()


def test_noalloc_spmv_refusals(gpu_ctx):
    g = gpu_ctx.rmat_graph(10, 8, 2)
    L = gpu_ctx.L
    x = DeviceArray(np.ones(g.N, np.float32))
    y = DeviceArray(np.zeros(g.N, np.float32))
    # no plan for this direction yet: building one would allocate
    rc = L.gbx_spmv_pull_noalloc(gpu_ctx.h, g.h, 1, 0, gbx.PULL_FWD, x.ptr, y.ptr)
    assert rc == -5 and b"no plan" in L.gbx_last_error(gpu_ctx.h)
    g.spmv(iterations=1, direction=gbx.PULL_FWD)
    assert L.gbx_spmv_pull_noalloc(gpu_ctx.h, g.h, 1, 0, gbx.PULL_FWD, None, y.ptr) == -1
    assert L.gbx_spmv_pull_noalloc(gpu_ctx.h, g.h, 1, 1, gbx.PULL_FWD, x.ptr, y.ptr) == -5  # unweighted graph
    assert L.gbx_spmv_pull_noalloc(gpu_ctx.h, g.h, 1, 0, gbx.PULL_FWD, x.ptr, y.ptr) == 0
    gpu_ctx.sync()
    x.free()
    y.free()
    g.free()
