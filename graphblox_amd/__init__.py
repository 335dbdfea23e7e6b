"""graphblox_amd -- MI355X (gfx950) backend for GraphBlox's pull PageRank / SpMV path.

The product is the native C ABI in ``libgbx_hip.so`` (declared in
``include/gbx.h``) and the reference-API adapter ``libgraphblox_hip.so``
(``include/graphblox_hip.h``).  This module is a thin ctypes mirror used by
tests and the benchmark: every call goes straight to the HIP library, and a
missing or unloadable library raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from typing import Optional

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(PKG, "libgbx_hip.so")
LIB_HOST = os.path.join(PKG, "libgraphblox_hip.so")

GBX_OK = 0
SORT_RADIX_SRC, SORT_RADIX_SRC_DST, SORT_COUNT_SRC, SORT_COUNT_SRC_DST = 0, 1, 2, 3
GRAPH_DEDUPE, GRAPH_WEIGHTED, GRAPH_SYMMETRIC = 1, 2, 4
PR_REF, PR_STOCHASTIC = 0, 1
PR_FIXED64, PR_FIXED32, PR_FIXED16, PR_FIXED8 = 2, 3, 4, 5
PR_QUANT32, PR_QUANT16, PR_QUANT8 = 6, 7, 8
PR_DATA_DRIVEN, PR_DD_PUSH, PR_DD_PULL_PUSH = 9, 10, 11
STOP_ACTIVE, STOP_L1, STOP_NONE = 0, 1, 2
F32, BF16 = 0, 1
PULL_IN, PULL_FWD = 0, 1
METHOD_AUTO, METHOD_PULL, METHOD_BINNED = 0, 1, 2


class GbxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gbx status {status}: {msg}")
        self.status = status


class PrParams(C.Structure):
    _fields_ = [("damp", C.c_float), ("max_iters", C.c_uint32), ("epsilon", C.c_double),
                ("mode", C.c_int), ("stop", C.c_int), ("precision", C.c_int),
                ("direction", C.c_int), ("weighted", C.c_int), ("method", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("bin_thresholds", C.c_uint32 * 7), ("hub_chunk", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("ppr_hub", C.c_uint32), ("ppr_chunk", C.c_uint32),
                ("ppr_blocks_per_cu", C.c_uint32), ("ppr_vectors_per_lane", C.c_uint32),
                ("binned_min_vertices", C.c_uint32), ("binned_chunk_edges", C.c_uint32),
                ("binned_bin_edges", C.c_uint32), ("binned_bin_rows", C.c_uint32),
                ("binned_chunk_rows", C.c_uint32), ("binned_hot_sources", C.c_uint32),
                ("binned_edge_values", C.c_uint32), ("reserved", C.c_uint32 * 1)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("bin_thresholds", "reserved")}
        d["bin_thresholds"] = list(self.bin_thresholds)
        return d


class Relabel(C.Structure):
    _fields_ = [("lmode", C.c_uint32 * 3), ("file_labels", C.c_void_p), ("num_file_labels", C.c_uint32)]


class PrReport(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("executed", C.c_uint32), ("active", C.c_uint32),
                ("converged", C.c_uint32), ("error_total", C.c_double), ("rank_sum", C.c_double),
                ("time_total_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_vp = C.c_void_p


def lib():
    """Load libgbx_hip.so (built by ``python -m graphblox_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_HIP):
        raise RuntimeError(f"{LIB_HIP} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_HIP)
    u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "gbx_init": (i32, [i32, C.POINTER(_vp)]),
        "gbx_free": (i32, [_vp]),
        "gbx_last_error": (C.c_char_p, [_vp]),
        "gbx_status_string": (C.c_char_p, [i32]),
        "gbx_sync": (i32, [_vp]),
        "gbx_stream": (_vp, [_vp]),
        "gbx_version": (C.c_char_p, []),
        "gbx_csr_build": (i32, [_vp, _vp, _vp, _vp, u32, u32, i32, i32, _vp, _vp, _vp, _vp, _vp]),
        "gbx_rank_sort": (i32, [_vp, _vp, u32, _vp]),
        "gbx_graph_create": (i32, [_vp, _vp, _vp, _vp, u32, u32, i32, C.c_uint, C.POINTER(_vp)]),
        "gbx_graph_create_relabel": (i32, [_vp, _vp, _vp, _vp, u32, u32, i32, C.c_uint, C.POINTER(Relabel), _vp,
                                           C.POINTER(_vp)]),
        "gbx_graph_create_rmat": (i32, [_vp, u32, u32, f64, f64, f64, u64, i32, u64, i32, C.c_uint,
                                        C.POINTER(_vp)]),
        "gbx_rmat_edges": (i32, [_vp, u32, u32, f64, f64, f64, u64, i32, i32, u64, _vp, _vp, _vp,
                                 C.POINTER(u64)]),
        "gbx_graph_from_csr": (i32, [_vp, u32, u32] + [_vp] * 10 + [C.POINTER(_vp)]),
        "gbx_graph_info": (i32, [_vp, C.POINTER(u32), C.POINTER(u32)]),
        "gbx_graph_export": (i32, [_vp, i32, _vp, _vp, _vp, _vp, _vp, _vp]),
        "gbx_graph_free": (i32, [_vp]),
        "gbx_pr_params_default": (None, [C.POINTER(PrParams)]),
        "gbx_pr_pull": (i32, [_vp, _vp, C.POINTER(PrParams), _vp, _vp, C.POINTER(PrReport)]),
        "gbx_pr_create": (i32, [_vp, _vp, C.POINTER(PrParams), C.POINTER(_vp)]),
        "gbx_pr_reset": (i32, [_vp, _vp]),
        "gbx_pr_launch": (i32, [_vp, u32]),
        "gbx_pr_launch_timed": (i32, [_vp, u32, C.POINTER(f64)]),
        "gbx_pr_read": (i32, [_vp, _vp, _vp]),
        "gbx_pr_free": (i32, [_vp]),
        "gbx_pr_plan_info": (i32, [_vp, C.POINTER(u32), _vp, _vp, _vp]),
        "gbx_spmv_pull": (i32, [_vp, _vp, u32, i32, i32, _vp, _vp, _vp, C.POINTER(f64)]),
        "gbx_spmv_fixed": (i32, [_vp, _vp, u32, i32, i32, i32, _vp, _vp, C.POINTER(f64)]),
        "gbx_spmv_pull_noalloc": (i32, [_vp, _vp, u32, i32, i32, _vp, _vp]),
        "gbx_gather_accumulate_u32": (i32, [_vp, _vp, i32, _vp, _vp]),
        "gbx_ppr_create": (i32, [_vp, _vp, C.POINTER(PrParams), u32, _vp, _vp, C.POINTER(_vp)]),
        "gbx_ppr_info": (i32, [_vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
        "gbx_pr_launch_graph": (i32, [_vp, u32]),
        "gbx_pr_adjoint": (i32, [_vp, _vp, C.POINTER(PrParams), _vp, _vp, _vp, _vp, _vp, _vp,
                                 C.POINTER(PrReport), C.POINTER(PrReport)]),
        "gbx_tuning_get": (i32, [_vp, C.POINTER(Tuning)]),
        "gbx_tuning_set": (i32, [_vp, C.POINTER(Tuning)]),
        "gbx_autotune": (i32, [_vp, _vp, C.POINTER(PrParams), u32, u32, C.POINTER(Tuning), C.POINTER(f64)]),
        "gbx_graph_batch_create": (i32, [_vp, u32, _vp, C.POINTER(_vp)]),
        "gbx_graph_batch_info": (i32, [_vp, C.POINTER(u32), _vp]),
        "gbx_edges_from_records": (i32, [_vp, _vp, u64, i32, i32, i32, C.POINTER(_vp)]),
        "gbx_edges_read_file": (i32, [_vp, C.c_char_p, i32, i32, i32, C.POINTER(_vp)]),
        "gbx_edges_info": (i32, [_vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_float)]),
        "gbx_edges_export": (i32, [_vp, _vp, _vp, _vp]),
        "gbx_graph_create_from_edges": (i32, [_vp, _vp, i32, C.c_uint, _vp, _vp, C.POINTER(_vp)]),
        "gbx_edges_free": (i32, [_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/gbx.h that this binding expects."""
    return [
        "gbx_init", "gbx_free", "gbx_last_error", "gbx_status_string", "gbx_sync", "gbx_stream",
        "gbx_version", "gbx_csr_build", "gbx_rank_sort", "gbx_graph_create", "gbx_graph_create_relabel",
        "gbx_graph_create_rmat",
        "gbx_rmat_edges", "gbx_graph_from_csr", "gbx_graph_info", "gbx_graph_export", "gbx_graph_free",
        "gbx_pr_params_default", "gbx_pr_pull", "gbx_pr_create", "gbx_pr_reset", "gbx_pr_launch",
        "gbx_pr_launch_timed", "gbx_pr_read", "gbx_pr_free", "gbx_pr_plan_info", "gbx_spmv_pull", "gbx_spmv_fixed",
        "gbx_ppr_create", "gbx_ppr_info", "gbx_pr_launch_graph", "gbx_graph_batch_create",
        "gbx_graph_batch_info", "gbx_tuning_get", "gbx_tuning_set", "gbx_autotune",
        "gbx_pr_adjoint", "gbx_spmv_pull_noalloc", "gbx_gather_accumulate_u32",
        "gbx_edges_from_records", "gbx_edges_read_file", "gbx_edges_info", "gbx_edges_export",
        "gbx_graph_create_from_edges", "gbx_edges_free",
    ]


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class Context:
    """A gbx_ctx on one HIP device."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = _vp()
        rc = self.L.gbx_init(device, C.byref(h))
        if rc != GBX_OK:
            raise GbxError(rc, self.L.gbx_last_error(None).decode())
        self.h = h
        # graphs and PageRank objects must be released before the context
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)
        return obj

    def check(self, rc: int):
        if rc != GBX_OK:
            raise GbxError(rc, self.L.gbx_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            for child in list(self._children):
                if isinstance(child, PageRank):
                    child.free()
            for child in list(self._children):
                child.free()
            self.L.gbx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.gbx_sync(self.h))

    def tuning(self) -> Tuning:
        t = Tuning()
        self.check(self.L.gbx_tuning_get(self.h, C.byref(t)))
        return t

    def set_tuning(self, t: Tuning):
        self.check(self.L.gbx_tuning_set(self.h, C.byref(t)))

    def autotune(self, graph, num_vectors=0, iters_per_trial=5, **kw):
        p = params(**kw)
        best = Tuning()
        ms = C.c_double()
        self.check(self.L.gbx_autotune(self.h, graph.h, C.byref(p), num_vectors, iters_per_trial, C.byref(best),
                                       C.byref(ms)))
        return best, ms.value

    # -- CSR build
    def csr_build(self, src, dst, w, num_vertices, sort_mode, move_weights=False):
        src, dst, w = _u32(src), _u32(dst), _f32(w)
        E = src.size
        out = dict(src=np.empty(E, np.uint32), dst=np.empty(E, np.uint32), w=np.empty(E, np.float32),
                   out_degree=np.empty(num_vertices, np.uint32), edges_idx=np.empty(num_vertices, np.uint32))
        self.check(self.L.gbx_csr_build(self.h, _ptr(src), _ptr(dst), _ptr(w), E, num_vertices, sort_mode,
                                        int(move_weights), _ptr(out["src"]), _ptr(out["dst"]), _ptr(out["w"]),
                                        _ptr(out["out_degree"]), _ptr(out["edges_idx"])))
        return out

    def rank_sort(self, ranks, labels=None):
        """radixSortEdgesByPageRank on the GPU: labels (default: the identity)
        stably sorted by the ascending fixed-point key of ranks."""
        ranks = np.ascontiguousarray(ranks, np.float32)
        lab = np.arange(ranks.size, dtype=np.uint32) if labels is None else _u32(labels).copy()
        self.check(self.L.gbx_rank_sort(self.h, _ptr(ranks), ranks.size, _ptr(lab)))
        return lab

    def graph(self, src, dst, w=None, num_vertices=0, sort_mode=SORT_RADIX_SRC_DST, flags=0, lmode=(0, 0, 0),
              file_labels=None) -> "Graph":
        """A device graph from an edge list; lmode: the relabel levels of
        graphCSRPreProcessingStep (graph.labels: new id of each vertex)."""
        src, dst, w = _u32(src), _u32(dst), _f32(w)
        h = _vp()
        if not any(lmode) and file_labels is None:
            self.check(self.L.gbx_graph_create(self.h, _ptr(src), _ptr(dst), _ptr(w), src.size, num_vertices,
                                               sort_mode, flags, C.byref(h)))
            return Graph(self, h)
        rl = Relabel()
        rl.lmode[:] = list(lmode) + [0] * (3 - len(lmode))
        fl = _u32(file_labels) if file_labels is not None else None
        rl.file_labels = _ptr(fl)
        rl.num_file_labels = fl.size if fl is not None else 0
        n = num_vertices or (int(max(src.max(), dst.max())) + 1 if src.size else 0)
        labels = np.zeros(max(n, 1), np.uint32)
        self.check(self.L.gbx_graph_create_relabel(self.h, _ptr(src), _ptr(dst), _ptr(w), src.size, num_vertices,
                                                   sort_mode, flags, C.byref(rl), _ptr(labels), C.byref(h)))
        g = Graph(self, h)
        g.labels = labels[:g.N]
        return g

    def graph_from_csr(self, fwd: dict, inv: dict) -> "Graph":
        """Upload CSR arrays in the reference layout (keys out_degree,
        edges_idx, src, dst[, w]) without re-sorting."""
        N, E = fwd["out_degree"].size, fwd["src"].size
        arrs = [_u32(fwd["out_degree"]), _u32(fwd["edges_idx"]), _u32(fwd["src"]), _u32(fwd["dst"]),
                _f32(fwd.get("w")), _u32(inv["out_degree"]), _u32(inv["edges_idx"]), _u32(inv["src"]),
                _u32(inv["dst"]), _f32(inv.get("w"))]
        h = _vp()
        self.check(self.L.gbx_graph_from_csr(self.h, N, E, *(_ptr(a) for a in arrs), C.byref(h)))
        return Graph(self, h)

    def graph_batch(self, graphs) -> "Graph":
        """Block-diagonal batch of independent graphs (copied)."""
        arr = (_vp * len(graphs))(*[g.h for g in graphs])
        h = _vp()
        self.check(self.L.gbx_graph_batch_create(self.h, len(graphs), C.cast(arr, _vp), C.byref(h)))
        g = Graph(self, h)
        g._batch = True
        return g

    def edges_read_file(self, path, weighted_records=False, inverse=False, symmetric=False) -> "Edges":
        """readEdgeListsbin on the GPU: the file's raw records to a device edge list."""
        h = _vp()
        self.check(self.L.gbx_edges_read_file(self.h, os.fsencode(path), int(weighted_records), int(inverse),
                                              int(symmetric), C.byref(h)))
        return Edges(self, h)

    def edges_from_records(self, records, weighted_records=False, inverse=False, symmetric=False) -> "Edges":
        """The same from records in host memory (bytes, or a numpy array of them)."""
        buf = np.frombuffer(records, np.uint8) if isinstance(records, (bytes, bytearray, memoryview)) else \
            np.ascontiguousarray(records).view(np.uint8).reshape(-1)
        h = _vp()
        self.check(self.L.gbx_edges_from_records(self.h, _ptr(buf) if buf.size else None, buf.size,
                                                 int(weighted_records), int(inverse), int(symmetric), C.byref(h)))
        return Edges(self, h)

    def rmat_graph(self, scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, weight_kind=0, weight_seed=0,
                   sort_mode=SORT_RADIX_SRC_DST, flags=0) -> "Graph":
        h = _vp()
        self.check(self.L.gbx_graph_create_rmat(self.h, scale, edge_factor, a, b, c, seed, weight_kind,
                                                weight_seed, sort_mode, flags, C.byref(h)))
        return Graph(self, h)

    def rmat_edges(self, scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, dedupe=True, weight_kind=0,
                   weight_seed=0):
        M = edge_factor << scale
        s = np.empty(M, np.uint32)
        d = np.empty(M, np.uint32)
        w = np.empty(M, np.float32)
        n = C.c_uint64()
        self.check(self.L.gbx_rmat_edges(self.h, scale, edge_factor, a, b, c, seed, int(dedupe), weight_kind,
                                         weight_seed, _ptr(s), _ptr(d), _ptr(w), C.byref(n)))
        k = n.value
        return s[:k].copy(), d[:k].copy(), w[:k].copy()


class Edges:
    """A device-resident edge list (gbx_edges): what readEdgeListsbin reads."""

    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h
        ctx._adopt(self)
        n, e, m = C.c_uint32(), C.c_uint32(), C.c_float()
        ctx.check(ctx.L.gbx_edges_info(h, C.byref(n), C.byref(e), C.byref(m)))
        self.N, self.E, self.max_weight = n.value, e.value, m.value

    def export(self):
        E = self.E
        s, d, w = np.empty(E, np.uint32), np.empty(E, np.uint32), np.empty(E, np.float32)
        self.ctx.check(self.ctx.L.gbx_edges_export(self.h, _ptr(s), _ptr(d), _ptr(w)))
        return s, d, w

    def graph(self, sort_mode=SORT_RADIX_SRC_DST, flags=0, lmode=(0, 0, 0), file_labels=None) -> "Graph":
        """gbx_graph_create_from_edges: the edge list is consumed."""
        rl = None
        labels = None
        if any(lmode) or file_labels is not None:
            rl = Relabel()
            rl.lmode[:] = list(lmode) + [0] * (3 - len(lmode))
            fl = _u32(file_labels) if file_labels is not None else None
            rl.file_labels = _ptr(fl)
            rl.num_file_labels = fl.size if fl is not None else 0
            labels = np.zeros(max(self.N, 1), np.uint32)
        h = _vp()
        try:
            self.ctx.check(self.ctx.L.gbx_graph_create_from_edges(self.ctx.h, self.h, sort_mode, flags,
                                                                  C.byref(rl) if rl is not None else None,
                                                                  _ptr(labels), C.byref(h)))
        finally:
            self.E = 0
        g = Graph(self.ctx, h)
        if labels is not None:
            g.labels = labels[:g.N]
        return g

    def free(self):
        if self.h:
            self.ctx.L.gbx_edges_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """A device-resident gbx_graph (forward + transpose CSR)."""

    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h
        ctx._adopt(self)
        n, e = C.c_uint32(), C.c_uint32()
        ctx.check(ctx.L.gbx_graph_info(h, C.byref(n), C.byref(e)))
        self.N, self.E = n.value, e.value

    def batch_offsets(self):
        n = C.c_uint32()
        self.ctx.check(self.ctx.L.gbx_graph_batch_info(self.h, C.byref(n), None))
        off = np.zeros(n.value + 1, np.uint32)
        self.ctx.check(self.ctx.L.gbx_graph_batch_info(self.h, C.byref(n), _ptr(off)))
        return off

    def export(self, inverse=False):
        N, E = self.N, self.E
        o = dict(out_degree=np.empty(N, np.uint32), in_degree=np.empty(N, np.uint32),
                 edges_idx=np.empty(N, np.uint32), src=np.empty(E, np.uint32), dst=np.empty(E, np.uint32),
                 w=np.empty(E, np.float32))
        self.ctx.check(self.ctx.L.gbx_graph_export(self.h, int(inverse), *(_ptr(o[k]) for k in
                                                   ("out_degree", "in_degree", "edges_idx", "src", "dst", "w"))))
        return o

    def free(self):
        if self.h:
            for child in list(self.ctx._children):
                if isinstance(child, PageRank) and child.g is self:
                    child.free()
            self.ctx.L.gbx_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def pagerank(self, iterations=20, epsilon=1e-4, mode=PR_REF, stop=STOP_ACTIVE, damp=0.85,
                 direction=PULL_IN, weighted=False, precision=F32, teleport=None, method=METHOD_AUTO):
        p = params(iterations, epsilon, mode, stop, damp, direction, weighted, precision, method)
        ranks = np.empty(self.N, np.float32)
        rep = PrReport()
        tp = _f32(teleport)
        self.ctx.check(self.ctx.L.gbx_pr_pull(self.ctx.h, self.h, C.byref(p), _ptr(tp), _ptr(ranks), C.byref(rep)))
        return ranks, rep.as_dict()

    def adjoint(self, loss_weights, teleport=None, iterations=300, epsilon=1e-7, damp=0.85, stop=STOP_L1,
                method=METHOD_AUTO):
        """Forward weighted stochastic PageRank p and its adjoint: returns
        p, lambda, dL/dW (forward-CSR edge order), dL/dt and both reports."""
        p = params(iterations, epsilon, PR_STOCHASTIC, stop, damp, PULL_IN, True, method=method)
        N, E = self.N, self.E
        out = dict(p=np.empty(N, np.float32), lam=np.empty(N, np.float32), grad_w=np.empty(E, np.float32),
                   grad_t=np.empty(N, np.float32))
        fr, ar = PrReport(), PrReport()
        g = _f32(loss_weights)
        tp = _f32(teleport)
        self.ctx.check(self.ctx.L.gbx_pr_adjoint(self.ctx.h, self.h, C.byref(p), _ptr(tp), _ptr(g), _ptr(out["p"]),
                                                 _ptr(out["lam"]), _ptr(out["grad_w"]), _ptr(out["grad_t"]),
                                                 C.byref(fr), C.byref(ar)))
        out["fwd_report"], out["adj_report"] = fr.as_dict(), ar.as_dict()
        return out

    def spmv(self, iterations=1, weighted=False, direction=PULL_IN, x=None):
        xo = np.empty(self.N, np.float32)
        yo = np.empty(self.N, np.float32)
        sec = C.c_double()
        xi = _f32(x)
        self.ctx.check(self.ctx.L.gbx_spmv_pull(self.ctx.h, self.h, iterations, int(weighted), direction, _ptr(xi),
                                                _ptr(xo), _ptr(yo), C.byref(sec)))
        return xo, yo, sec.value

    def spmv_noalloc(self, x_ptr: int, y_ptr: int, iterations=1, weighted=False, direction=PULL_IN):
        """gbx_spmv_pull_noalloc on device arrays x, y [N] (vertex order):
        launched on the context's stream, not waited for."""
        self.ctx.check(self.ctx.L.gbx_spmv_pull_noalloc(self.ctx.h, self.h, iterations, int(weighted), direction,
                                                         x_ptr, y_ptr))

    def gather_accumulate_u32(self, aux1, aux2=None, direction=PULL_IN):
        """The FPGA overlay's Single.PR integer pull on u32[2N] buffers:
        aux2[N + v] = sum of aux1[N + u] over row v (mod 2^32)."""
        a1 = _u32(aux1)
        if a1.size != 2 * self.N:
            raise ValueError("aux1 must hold 2N words")
        a2 = np.zeros(2 * self.N, np.uint32) if aux2 is None else _u32(aux2).copy()
        self.ctx.check(self.ctx.L.gbx_gather_accumulate_u32(self.ctx.h, self.h, direction, _ptr(a1), _ptr(a2)))
        return a2

    def spmv_fixed(self, bits, iterations=1, weighted=False, direction=PULL_IN):
        """The reference's fixed-point SpMV variants (bits 32: pull, 64: push)."""
        xo = np.empty(self.N, np.float32)
        yo = np.empty(self.N, np.float32)
        sec = C.c_double()
        self.ctx.check(self.ctx.L.gbx_spmv_fixed(self.ctx.h, self.h, iterations, int(weighted), direction, bits,
                                                 _ptr(xo), _ptr(yo), C.byref(sec)))
        return xo, yo, sec.value

    def session(self, **kw) -> "PageRank":
        return PageRank(self, params(**kw))

    def ppr_session(self, num_vectors, seed_sets=None, iterations=100, epsilon=1e-7, stop=STOP_L1, damp=0.85,
                    direction=PULL_IN) -> "PageRank":
        """Batched personalised PageRank: seed_sets is a list of vertex-id
        arrays (one per vector) or None for uniform teleports."""
        p = params(iterations, epsilon, PR_STOCHASTIC, stop, damp, direction)
        return PageRank(self, p, num_vectors=num_vectors, seed_sets=seed_sets)


def params(iterations=20, epsilon=1e-4, mode=PR_REF, stop=STOP_ACTIVE, damp=0.85, direction=PULL_IN,
           weighted=False, precision=F32, method=METHOD_AUTO) -> PrParams:
    p = PrParams()
    lib().gbx_pr_params_default(C.byref(p))
    p.damp = damp
    p.max_iters = iterations
    p.epsilon = epsilon
    p.mode = mode
    p.stop = stop
    p.precision = precision
    p.direction = direction
    p.weighted = int(weighted)
    p.method = method
    return p


class PageRank:
    """Step-wise gbx_pr object (create / reset / launch / read)."""

    def __init__(self, graph: Graph, p: PrParams, num_vectors: int = 0, seed_sets=None):
        self.g = graph
        self.ctx = graph.ctx
        self.k = num_vectors
        h = _vp()
        if num_vectors:
            if seed_sets is not None:
                sizes = [len(s) for s in seed_sets]
                seed_ptr = np.zeros(num_vectors + 1, np.uint32)
                seed_ptr[1:] = np.cumsum(sizes)
                seeds = _u32(np.concatenate([np.asarray(s, np.uint32) for s in seed_sets]))
                self.ctx.check(self.ctx.L.gbx_ppr_create(self.ctx.h, graph.h, C.byref(p), num_vectors,
                                                         _ptr(seed_ptr), _ptr(seeds), C.byref(h)))
            else:
                self.ctx.check(self.ctx.L.gbx_ppr_create(self.ctx.h, graph.h, C.byref(p), num_vectors, None,
                                                         None, C.byref(h)))
        else:
            self.ctx.check(self.ctx.L.gbx_pr_create(self.ctx.h, graph.h, C.byref(p), C.byref(h)))
        self.h = h
        self.ctx._adopt(self)
        self.nreports = 0 if num_vectors else len(graph.batch_offsets()) - 1
        if self.nreports == 1 and graph.N and not self._is_batch(graph):
            self.nreports = 0

    @staticmethod
    def _is_batch(graph):
        return getattr(graph, "_batch", False)

    def reset(self, teleport=None):
        tp = _f32(teleport)
        self.ctx.check(self.ctx.L.gbx_pr_reset(self.h, _ptr(tp)))

    def launch(self, n):
        self.ctx.check(self.ctx.L.gbx_pr_launch(self.h, n))

    def launch_graph(self, n):
        self.ctx.check(self.ctx.L.gbx_pr_launch_graph(self.h, n))

    def launch_timed(self, n) -> float:
        ms = C.c_double()
        self.ctx.check(self.ctx.L.gbx_pr_launch_timed(self.h, n, C.byref(ms)))
        return ms.value

    def read(self):
        if self.nreports > 1:
            ranks = np.empty(self.g.N, np.float32)
            reps = (PrReport * self.nreports)()
            self.ctx.check(self.ctx.L.gbx_pr_read(self.h, _ptr(ranks), C.cast(reps, _vp)))
            return ranks, [r.as_dict() for r in reps]
        if self.k:
            ranks = np.empty((self.g.N, self.k), np.float32)
            reps = (PrReport * self.k)()
            self.ctx.check(self.ctx.L.gbx_pr_read(self.h, _ptr(ranks), C.cast(reps, _vp)))
            return ranks, [r.as_dict() for r in reps]
        ranks = np.empty(self.g.N, np.float32)
        rep = PrReport()
        self.ctx.check(self.ctx.L.gbx_pr_read(self.h, _ptr(ranks), C.cast(C.pointer(rep), _vp)))
        return ranks, rep.as_dict()

    def ppr_info(self):
        k, v, h, c = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.ctx.check(self.ctx.L.gbx_ppr_info(self.h, C.byref(k), C.byref(v), C.byref(h), C.byref(c)))
        return dict(num_vectors=k.value, vector_width=v.value, hub_rows=h.value, hub_chunks=c.value)

    def plan_info(self):
        n = C.c_uint32()
        lanes = np.zeros(16, np.uint32)
        rows = np.zeros(16, np.uint32)
        edges = np.zeros(16, np.uint64)
        self.ctx.check(self.ctx.L.gbx_pr_plan_info(self.h, C.byref(n), _ptr(lanes), _ptr(rows), _ptr(edges)))
        k = n.value
        return [dict(lanes=int(lanes[i]), rows=int(rows[i]), edges=int(edges[i])) for i in range(k)]

    def free(self):
        if self.h:
            self.ctx.L.gbx_pr_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
