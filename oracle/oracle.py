"""CPU oracle for the pull PageRank / SpMV path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only as the checker or as the
timed CPU baseline.  The product package (``graphblox_amd``) never imports it.

It wraps two native libraries:

* ``oracle/liboracle.so`` -- the clean-room C/OpenMP restatement in
  ``gbx_oracle.c`` (each function there cites the reference lines it follows);
* ``oracle/_ref/libgbxref_d<D>w<W>.so`` -- the GraphBlox reference's own host
  sources compiled by ``oracle/Makefile`` (present only where the reference
  sources were available at build time; optional everywhere).

and adds numpy helpers: the ``.bin``/``.wbin`` reader with the reference's
``N = max id + 1`` rule (01_host/src/structures/edgeList.c:450-706), the CSR
text-fixture parser (01_host/src/structures/graphCSR.c:529-635 writes them),
and dense FP64 closed forms used to pin the iterative oracles on tiny graphs.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

_lib = None
_ref = {}


def lib():
    """The clean-room oracle library (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_count_sort.argtypes = [_u32p, _u32p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int]
        L.orc_map_vertices.argtypes = [_u32p, C.c_uint32, C.c_uint32, _u32p, _u32p]
        L.orc_remove_duplicates.argtypes = [_u32p, _u32p, _f32p, C.c_uint32]
        L.orc_remove_duplicates.restype = C.c_uint32
        L.orc_pr_pull_f32.argtypes = [C.c_uint32, _u32p, _u32p, _u32p, _u32p, C.c_uint32, C.c_double,
                                      C.c_int, _f32p, C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                      C.c_void_p]
        L.orc_spmv_pull_f32.argtypes = [C.c_uint32, _u32p, _u32p, _u32p, _u32p, C.c_void_p, C.c_uint32,
                                        C.c_int, _f32p, _f32p, C.c_void_p]
        L.orc_pr_f64.argtypes = [C.c_uint32, _u32p, _u64p, _u32p, C.c_void_p, C.c_void_p, C.c_double,
                                 C.c_int, C.c_uint32, C.c_double, _f64p, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_double)]
        L.orc_pr_adjoint_f64.argtypes = [C.c_uint32, _u64p, _u32p, _f64p, C.c_void_p, _f64p, C.c_double,
                                         C.c_uint32, C.c_double, _f64p, _f64p, _f64p, _f64p]
        L.orc_pr_loss_f64.argtypes = [C.c_uint32, _u64p, _u32p, _f64p, C.c_void_p, _f64p, C.c_double,
                                      C.c_void_p, C.c_uint32, C.c_double, _f64p]
        L.orc_pr_loss_f64.restype = C.c_double
        L.orc_pr_loss_diff.argtypes = [C.c_uint32, _u64p, _u32p, _f64p, _f64p, C.c_void_p, _f64p, C.c_double,
                                       C.c_void_p, C.c_uint32]
        L.orc_pr_loss_diff.restype = C.c_double
        L.orc_pr_pull_fixed.argtypes = [C.c_uint32, _u32p, _u64p, _u32p, C.c_int, C.c_uint32, C.c_double, _f32p,
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.orc_pr_pull_quant.argtypes = L.orc_pr_pull_fixed.argtypes
        L.orc_pr_pull_dd.argtypes = [C.c_uint32, _u32p, _u64p, _u32p, _u64p, _u32p, C.c_uint32, C.c_double, _f32p,
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.orc_pr_dd_push.argtypes = [C.c_uint32, _u32p, _u64p, _u32p, _u64p, _u32p, C.c_int, C.c_uint32, C.c_double,
                                     _f32p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.orc_rmat_generate.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, _u32p, _u32p]
        L.orc_rmat_permute.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        L.orc_rmat_permute.restype = C.c_uint32
        L.orc_dedupe_keep_first.argtypes = [_u32p, _u32p, C.c_void_p, C.c_uint64]
        L.orc_dedupe_keep_first.restype = C.c_uint64
        _lib = L
    return _lib


def ref_available(directed: int = 1, weighted: int = 0) -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libgbxref_d{directed}w{weighted}.so"))


def ref(directed: int = 1, weighted: int = 0):
    """The reference's own code compiled from its sources (optional)."""
    key = (directed, weighted)
    if key not in _ref:
        path = os.path.join(HERE, "_ref", f"libgbxref_d{directed}w{weighted}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (needs the reference sources; `make -C oracle ref`)")
        L = C.CDLL(path)
        L.refd_build.argtypes = [C.c_char_p, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int]
        L.refd_build.restype = C.c_void_p
        L.refd_build_relabel.argtypes = [C.c_char_p, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_char_p]
        L.refd_build_relabel.restype = C.c_void_p
        L.refd_labels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.refd_num_vertices.argtypes = [C.c_void_p]
        L.refd_num_vertices.restype = C.c_uint32
        L.refd_num_edges.argtypes = [C.c_void_p]
        L.refd_num_edges.restype = C.c_uint32
        L.refd_export.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 6
        L.refd_free.argtypes = [C.c_void_p]
        L.refd_pagerank.argtypes = [C.c_void_p, C.c_uint32, C.c_double, C.c_int, _f32p,
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.refd_pagerank_variant.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_double, C.c_int, _f32p,
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.refd_spmv.argtypes = [C.c_void_p, C.c_uint32, C.c_int, _f32p, _f32p, C.POINTER(C.c_double)]
        L.refd_spmv_variant.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_int, _f32p, _f32p]
        if directed:
            L.refd_pagerank_on_csr.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p, _u32p,
                                               C.c_uint32, C.c_double, C.c_int, _f32p,
                                               C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double)]
        _ref[key] = L
    return _ref[key]


# --------------------------------------------------------------------------- I/O

def read_bin(path: str, weighted: bool = False):
    """Edge list reader: records of (u32 src, u32 dst[, f32 w]); N = max id + 1
    (01_host/src/structures/edgeList.c:495-519, 671-679)."""
    raw = np.fromfile(path, dtype=np.uint32)
    rec = 3 if weighted else 2
    E = raw.size // rec
    raw = raw[: E * rec].reshape(E, rec)
    src = np.ascontiguousarray(raw[:, 0])
    dst = np.ascontiguousarray(raw[:, 1])
    w = np.ascontiguousarray(raw[:, 2]).view(np.float32) if weighted else np.ones(E, np.float32)
    N = int(max(src.max(initial=0), dst.max(initial=0))) + 1 if E else 1
    return src, dst, w, N


def write_bin(path: str, src, dst, w=None):
    if w is None:
        rec = np.stack([src.astype(np.uint32), dst.astype(np.uint32)], axis=1)
    else:
        rec = np.stack([src.astype(np.uint32), dst.astype(np.uint32),
                        np.asarray(w, np.float32).view(np.uint32)], axis=1)
    np.ascontiguousarray(rec).tofile(path)


def read_csr_text_fixture(prefix: str):
    """Parse the reference's CSR text dump (graphCSR.c:587-604): .out_degree
    and .edges_array_src start with a count line; the rest do not."""
    def col(ext, header):
        vals = np.loadtxt(prefix + ext, dtype=np.int64, ndmin=1)
        return (int(vals[0]), vals[1:].astype(np.uint32)) if header else (None, vals.astype(np.uint32))
    N, out_degree = col(".out_degree", True)
    E, src = col(".edges_array_src", True)
    _, dst = col(".edges_array_dest", False)
    _, idx = col(".edges_idx", False)
    _, ind = col(".in_degree", False)
    return dict(N=N, E=E, out_degree=out_degree, edges_idx=idx, in_degree=ind, src=src, dst=dst)


# --------------------------------------------------------------------------- CSR build

def count_sort(src, dst, w, N, mode, move_weights=False):
    s = np.ascontiguousarray(src, np.uint32).copy()
    d = np.ascontiguousarray(dst, np.uint32).copy()
    ww = np.ascontiguousarray(w, np.float32).copy() if w is not None else np.ones(s.size, np.float32)
    rc = lib().orc_count_sort(s, d, ww.ctypes.data, s.size, N, mode, int(move_weights))
    assert rc == 0
    return s, d, ww


def map_vertices(sorted_src, N):
    od = np.zeros(N, np.uint32)
    idx = np.zeros(N, np.uint32)
    lib().orc_map_vertices(np.ascontiguousarray(sorted_src, np.uint32), sorted_src.size, N, od, idx)
    return od, idx


@dataclass
class OracleCSR:
    N: int
    E: int
    out_degree: np.ndarray
    in_degree: np.ndarray
    edges_idx: np.ndarray
    src: np.ndarray
    dst: np.ndarray
    w: np.ndarray


def build_graph(src, dst, w, N, mode=1, dedupe=False, move_weights=False):
    """Forward CSR then transpose, in graphCSRPreProcessingStep's order
    (01_host/src/structures/graphCSR.c:243-327 with DIRECTED=1)."""
    if dedupe:
        mode = 1  # GraphBlox.c:358-359
    s, d, ww = count_sort(src, dst, w, N, mode, move_weights)
    if dedupe:
        E = lib().orc_remove_duplicates(s, d, ww, s.size)
        s, d, ww = s[:E].copy(), d[:E].copy(), ww[:E].copy()
    od, idx = map_vertices(s, N)
    # transpose: swap (edgeList.c:799-801) and run the same sort
    ts, td, tw = count_sort(d, s, ww, N, mode, move_weights)
    iod, iidx = map_vertices(ts, N)
    fwd = OracleCSR(N, s.size, od, iod.copy(), idx, s, d, ww)
    inv = OracleCSR(N, s.size, iod, od.copy(), iidx, ts, td, tw)
    return fwd, inv


def row_ptr_from(csr: OracleCSR) -> np.ndarray:
    rp = np.zeros(csr.N + 1, np.uint64)
    np.cumsum(csr.out_degree, out=rp[1:])
    return rp


# --------------------------------------------------------------------------- algorithms

def pr_pull_f32(fwd: OracleCSR, inv: OracleCSR, iterations=20, epsilon=0.0, nthreads=1,
                timings=False):
    """Reference-semantics FP32 pull PageRank (pageRank.c:245-368)."""
    N = fwd.N
    ranks = np.zeros(N, np.float32)
    it = C.c_uint32()
    err = C.c_double()
    ts = np.zeros(max(iterations, 1), np.float64)
    lib().orc_pr_pull_f32(N, fwd.out_degree, inv.out_degree, inv.edges_idx, inv.dst, iterations, epsilon,
                          nthreads, ranks, C.byref(it), C.byref(err), ts.ctypes.data)
    out = dict(ranks=ranks, iterations=it.value, error_total=err.value)
    if timings:
        out["iter_seconds"] = ts[: max(it.value, 1)]
    return out


def spmv_pull_f32(fwd: OracleCSR, inv: OracleCSR, iterations=1, weights=None, nthreads=1):
    N = fwd.N
    x = np.zeros(N, np.float32)
    y = np.zeros(N, np.float32)
    wp = np.ascontiguousarray(weights, np.float32) if weights is not None else None
    lib().orc_spmv_pull_f32(N, fwd.out_degree, inv.out_degree, inv.edges_idx, inv.dst,
                            wp.ctypes.data if wp is not None else None, iterations, nthreads, x, y, None)
    return x, y


def pr_f64(N, outdeg, row_ptr, col, mode=0, damp=0.85, iterations=20, epsilon=0.0,
           weights=None, teleport=None):
    """FP64 PageRank oracle (mode 0: reference semantics, 1: stochastic)."""
    r = np.zeros(N, np.float64)
    it = C.c_uint32()
    err = C.c_double()
    wp = np.ascontiguousarray(weights, np.float64) if weights is not None else None
    tp = np.ascontiguousarray(teleport, np.float64) if teleport is not None else None
    lib().orc_pr_f64(N, np.ascontiguousarray(outdeg, np.uint32), np.ascontiguousarray(row_ptr, np.uint64),
                     np.ascontiguousarray(col, np.uint32), wp.ctypes.data if wp is not None else None,
                     tp.ctypes.data if tp is not None else None, damp, mode, iterations, epsilon, r,
                     C.byref(it), C.byref(err))
    return r, it.value, err.value


def pr_pull_fixed(N, outdeg, row_ptr, col, bits, iterations=20, epsilon=0.0):
    """The reference's fixed-point pull PageRank (64/32/16/8-bit)."""
    r = np.zeros(N, np.float32)
    it, err = C.c_uint32(), C.c_double()
    rc = lib().orc_pr_pull_fixed(N, np.ascontiguousarray(outdeg, np.uint32), np.ascontiguousarray(row_ptr, np.uint64),
                                 np.ascontiguousarray(col, np.uint32), bits, iterations, epsilon, r, C.byref(it),
                                 C.byref(err))
    if rc:
        raise ValueError(f"bits {bits}")
    return r, it.value, err.value


def pr_pull_dd(N, outdeg, row_ptr, col, fwd_row_ptr, fwd_col, iterations=20, epsilon=0.0):
    """The reference's data-driven pull PageRank (worklist, conditional update)."""
    r = np.zeros(N, np.float32)
    it, err = C.c_uint32(), C.c_double()
    rc = lib().orc_pr_pull_dd(N, np.ascontiguousarray(outdeg, np.uint32), np.ascontiguousarray(row_ptr, np.uint64),
                              np.ascontiguousarray(col, np.uint32), np.ascontiguousarray(fwd_row_ptr, np.uint64),
                              np.ascontiguousarray(fwd_col, np.uint32), iterations, epsilon, r, C.byref(it),
                              C.byref(err))
    if rc:
        raise MemoryError("orc_pr_pull_dd")
    return r, it.value, err.value


def pr_dd_push(N, outdeg, row_ptr, col, fwd_row_ptr, fwd_col, pull_push=False, iterations=20, epsilon=0.0):
    """The reference's data-driven push (or pull-push) PageRank, sequential."""
    r = np.zeros(N, np.float32)
    it, err = C.c_uint32(), C.c_double()
    rc = lib().orc_pr_dd_push(N, np.ascontiguousarray(outdeg, np.uint32), np.ascontiguousarray(row_ptr, np.uint64),
                              np.ascontiguousarray(col, np.uint32), np.ascontiguousarray(fwd_row_ptr, np.uint64),
                              np.ascontiguousarray(fwd_col, np.uint32), int(pull_push), iterations, epsilon, r,
                              C.byref(it), C.byref(err))
    if rc:
        raise MemoryError("orc_pr_dd_push")
    return r, it.value, err.value


def dd_push_limit(N, outdeg, row_ptr, col, damp=0.85, tol=1e-15, max_iters=5000):
    """FP64 limit of the reference's data-driven push PageRank
    (pageRank.c:1872-2027), normalised by N: ranks start at 1-d with residual
    res0 = (1-d) d P^T 1 (self loops included, pageRank.c:1908-1923); every
    later push moves d * res / outdeg to each out-neighbour except the vertex
    itself (its own residual is zeroed after the push loop), so
    r = (1-d) 1 + (I - d P'^T)^-1 res0 with P' the transition matrix without
    self loops (out-degrees unchanged)."""
    od = np.asarray(outdeg, np.int64)
    rp = np.asarray(row_ptr, np.int64)
    c = np.asarray(col, np.int64) & 0x3FFFFFFF
    rows = np.repeat(np.arange(N), np.diff(rp))
    inv_od = np.where(od > 0, 1.0 / np.maximum(od, 1), 0.0)
    res0 = (1.0 - damp) * damp * np.bincount(rows, weights=inv_od[c], minlength=N)
    keep = rows != c
    kr, kc = rows[keep], c[keep]
    y = res0.copy()
    for _ in range(max_iters):
        y2 = res0 + damp * np.bincount(kr, weights=y[kc] * inv_od[kc], minlength=N)
        done = np.abs(y2 - y).sum() < tol * max(1.0, np.abs(y2).sum())
        y = y2
        if done:
            break
    return ((1.0 - damp) + y) / N


def mt19937_uint32(seed, count):
    """MT19937 (Matsumoto & Nishimura 1998) with the 1999 reference code's
    seeding (each state word from two steps of x <- 69069 x + 1, upper 16 bits
    each), as mt19937.c:46-57 initialises it; `count` tempered 32-bit outputs."""
    n, m = 624, 397
    mt = [0] * n
    x = seed & 0xFFFFFFFF
    for i in range(n):
        hi = x & 0xFFFF0000
        x = (69069 * x + 1) & 0xFFFFFFFF
        mt[i] = hi | ((x & 0xFFFF0000) >> 16)
        x = (69069 * x + 1) & 0xFFFFFFFF
    out, idx = [], n
    for _ in range(count):
        if idx >= n:
            for k in range(n):
                y = (mt[k] & 0x80000000) | (mt[(k + 1) % n] & 0x7FFFFFFF)
                mt[k] = mt[(k + m) % n] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y & 0xFFFFFFFF)
    return out


def relabel_labels(src, dst, N, avg_degree, lmode, file_labels=None):
    """New id of every vertex for one relabel level of reorderGraphProcess
    (reorder.c:1556-1616) on an edge list, one OpenMP thread:
      degrees (reorder.c:1483-1552): out (lmode 1, 4, 6, 8, 12, 13 and any
        other), in (2, 5, 7, 9), in + out (3), MT19937(27491095) % N (10);
      degree order (1, 2, 3, 10 and other codes, reorder.c:296-385): by
        degree descending, equal degrees by id descending;
      DBG and Corder (4, 5, 12, 13; Corder runs the DBG list function,
        reorder.c:574-624): 11 buckets, thresholds t0 = avg_degree / 2 (1 when
        avg_degree <= 1) doubling, the last unbounded; bucket = first with
        degree <= threshold; buckets from high to low, ids ascending within;
      HUBSort (6, 7): hubs (degree > t0) by degree descending, equal degrees by
        id descending, then the others by id ascending (reorder.c:814-982);
      HUBCluster (8, 9): hubs (degree > avg_degree, reorder.c:1003) then the
        others, ids ascending within each;
      file (11): the given labels."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    if lmode == 11:
        return np.asarray(file_labels, np.uint32)[:N].copy()
    if lmode == 10:
        deg = np.array(mt19937_uint32(27491095, N), np.uint64) % np.uint64(N)
    else:
        if lmode in (2, 5, 7, 9):
            deg = np.bincount(dst, minlength=N)
        elif lmode == 3:
            deg = np.bincount(dst, minlength=N) + np.bincount(src, minlength=N)
        else:
            deg = np.bincount(src, minlength=N)
    deg = np.asarray(deg, np.int64)
    ids = np.arange(N, dtype=np.int64)
    t0 = 1 if avg_degree <= 1 else avg_degree // 2
    if lmode in (4, 5, 12, 13, 8, 9):
        nb = 2 if lmode in (8, 9) else 11
        th = [avg_degree] if lmode in (8, 9) else [t0 * (1 << i) for i in range(nb - 1)]
        th = th + [0xFFFFFFFF]
        bucket = np.searchsorted(np.array(th, np.int64), deg, side="left")
        order = ids[np.lexsort((ids, -bucket))]
    elif lmode in (6, 7):
        hub = deg > t0
        h = ids[hub]
        hubs = h[np.lexsort((-h, -deg[hub]))]
        order = np.concatenate([hubs, ids[~hub]])
    else:
        order = ids[np.lexsort((-ids, -deg))]
    labels = np.empty(N, np.uint32)
    labels[order] = np.arange(N, dtype=np.uint32)
    return labels


def _x86_to_u64(x):
    """(uint64_t) of non-negative doubles below 2^64, as x86-64 code converts them."""
    x = np.asarray(x, np.float64)
    big = x >= 2.0**63
    out = np.where(big, x - 2.0**63, x).astype(np.int64).astype(np.uint64)
    return np.where(big, out + np.uint64(1 << 63), out)


def spmv_fixed(N, outdeg, row_ptr, col, w, bits, iterations=1):
    """The reference's fixed-point SpMV (SPMV.c:313-590): x = 1/outdeg (0.001f
    when 0) as float.  bits 32 (SPMVPullFixedPointGraphCSR): x32 =
    FloatToFixed32(x), weight = (float)FloatToFixed32(w or 0.0001f), term =
    ((u64)weight * x32) >> 16; bits 64 (SPMVPushFixedPointGraphCSR, the same
    integer terms pushed): x64 = DoubleToFixed64(x), w64 = DoubleToFixed64(w or
    0.0001f), term = ((u128)w64 * x64) >> 32 mod 2^64.  y (u64) accumulates
    `iterations` passes; out: float y = Fixed32ToFloat / (float)Fixed64ToDouble.
    row_ptr/col/w: the rows the pull runs over (in-CSR for DIRECTED)."""
    od = np.asarray(outdeg, np.uint32)
    x = np.where(od > 0, np.float32(1.0) / np.maximum(od, 1).astype(np.float32), np.float32(0.001)).astype(np.float32)
    rp = np.asarray(row_ptr, np.int64)
    c = np.asarray(col, np.int64) & 0x3FFFFFFF
    rows = np.repeat(np.arange(N), np.diff(rp))
    if bits == 32:
        def f2x32(v):
            t = (np.asarray(v, np.float32) * np.float32(65536.0) + np.float32(0.5)).astype(np.float32)
            return t.astype(np.int64).astype(np.uint32)
        x32 = f2x32(x).astype(np.uint64)
        wq = (f2x32(w) if w is not None else np.full(c.size, f2x32(np.float32(0.0001)), np.uint32))
        wq = wq.astype(np.float32).astype(np.uint64)
        term = (wq * x32[c]) >> np.uint64(16)
    else:
        def d2x64(v):
            return _x86_to_u64(np.asarray(v, np.float32).astype(np.float64) * 4294967296.0 + 0.5)
        x64 = d2x64(x)
        wq = d2x64(w) if w is not None else np.full(c.size, d2x64(np.float32(0.0001)), np.uint64)
        a, b = wq, x64[c]
        a_hi, a_lo = a >> np.uint64(32), a & np.uint64(0xFFFFFFFF)
        b_hi, b_lo = b >> np.uint64(32), b & np.uint64(0xFFFFFFFF)
        # ((a * b) >> 32) mod 2^64 from 32-bit limbs
        lo = a_lo * b_lo
        mid = a_hi * b_lo + a_lo * b_hi + (lo >> np.uint64(32))  # may wrap: only its low 64 bits matter below
        term = (a_hi * b_hi << np.uint64(32)) + mid
    y = np.zeros(N, np.uint64)
    np.add.at(y, rows, term)
    y = y * np.uint64(iterations)
    if bits == 32:
        out = y.astype(np.float32) / np.float32(65536.0)
    else:
        out = (y.astype(np.float64) / 4294967296.0).astype(np.float32)
    return x, out.astype(np.float32)


def pr_pull_quant(N, outdeg, row_ptr, col, bits, iterations=20, epsilon=0.0):
    """The reference's quantized pull PageRank (32/16/8-bit, min/max scale)."""
    r = np.zeros(N, np.float32)
    it, err = C.c_uint32(), C.c_double()
    rc = lib().orc_pr_pull_quant(N, np.ascontiguousarray(outdeg, np.uint32), np.ascontiguousarray(row_ptr, np.uint64),
                                 np.ascontiguousarray(col, np.uint32), bits, iterations, epsilon, r, C.byref(it),
                                 C.byref(err))
    if rc:
        raise ValueError(f"bits {bits}")
    return r, it.value, err.value


def adjoint_f64(N, row_ptr, col, wraw, g, teleport=None, damp=0.85, iterations=1000, tol=1e-14):
    p = np.zeros(N)
    lam = np.zeros(N)
    gw = np.zeros(int(row_ptr[-1]))
    gt = np.zeros(N)
    tp = np.ascontiguousarray(teleport, np.float64) if teleport is not None else None
    lib().orc_pr_adjoint_f64(N, np.ascontiguousarray(row_ptr, np.uint64), np.ascontiguousarray(col, np.uint32),
                             np.ascontiguousarray(wraw, np.float64), tp.ctypes.data if tp is not None else None,
                             np.ascontiguousarray(g, np.float64), damp, iterations, tol, p, lam, gw, gt)
    return p, lam, gw, gt


def loss_f64(N, row_ptr, col, wraw, g, teleport=None, damp=0.85, p0=None, iterations=1000, tol=0.0):
    """L = g.p at the weighted stochastic fixed point (FP64), and p."""
    p = np.zeros(N)
    tp = np.ascontiguousarray(teleport, np.float64) if teleport is not None else None
    p0c = np.ascontiguousarray(p0, np.float64) if p0 is not None else None
    L = lib().orc_pr_loss_f64(N, np.ascontiguousarray(row_ptr, np.uint64), np.ascontiguousarray(col, np.uint32),
                              np.ascontiguousarray(wraw, np.float64), tp.ctypes.data if tp is not None else None,
                              np.ascontiguousarray(g, np.float64), damp,
                              p0c.ctypes.data if p0c is not None else None, iterations, tol, p)
    return L, p


def loss_diff(N, row_ptr, col, w_plus, w_minus, g, teleport=None, damp=0.85, p0=None, iterations=100):
    """L(w_plus) - L(w_minus), both in extended precision from p0."""
    tp = np.ascontiguousarray(teleport, np.float64) if teleport is not None else None
    p0c = np.ascontiguousarray(p0, np.float64) if p0 is not None else None
    return lib().orc_pr_loss_diff(N, np.ascontiguousarray(row_ptr, np.uint64), np.ascontiguousarray(col, np.uint32),
                                  np.ascontiguousarray(w_plus, np.float64), np.ascontiguousarray(w_minus, np.float64),
                                  tp.ctypes.data if tp is not None else None, np.ascontiguousarray(g, np.float64),
                                  damp, p0c.ctypes.data if p0c is not None else None, iterations)


def fd_edge_grads(N, row_ptr, col, wraw, g, edges, p0, h=1e-3, teleport=None, damp=0.85, iterations=100):
    """FP64 central finite differences dL/dW at in-CSR edge positions."""
    out = np.empty(len(edges))
    for k, e in enumerate(edges):
        wp = np.array(wraw, np.float64)
        wm = wp.copy()
        wp[e] *= 1.0 + h
        wm[e] *= 1.0 - h
        out[k] = loss_diff(N, row_ptr, col, wp, wm, g, teleport, damp, p0, iterations) / (wp[e] - wm[e])
    return out


def weighted_rmat_workload(scale, edge_factor=16, seed=20261203):
    """The differentiable-PageRank workload: seeded R-MAT (permuted ids,
    duplicates removed), FP32 edge weights lognormal(0, 1) in edge-list order
    and loss weights g ~ N(0, 1) per vertex.  Returns src, dst, w, N, g."""
    s, d, N = rmat_graph(scale, edge_factor, seed + scale)
    w = np.random.default_rng(seed + 100 + scale).lognormal(0.0, 1.0, s.size).astype(np.float32)
    g = np.random.default_rng(seed + 200 + scale).standard_normal(N)
    return s, d, w, N, g


def rmat(scale, edge_factor, seed, a=0.57, b=0.19, c=0.19):
    M = edge_factor << scale
    src = np.zeros(M, np.uint32)
    dst = np.zeros(M, np.uint32)
    lib().orc_rmat_generate(scale, edge_factor, a, b, c, seed, src, dst)
    return src, dst


def rmat_permute(x: int, scale: int, seed: int) -> int:
    return int(lib().orc_rmat_permute(x, scale, seed))


def dedupe_keep_first(src, dst, w=None):
    s = np.ascontiguousarray(src, np.uint32).copy()
    d = np.ascontiguousarray(dst, np.uint32).copy()
    ww = np.ascontiguousarray(w, np.float32).copy() if w is not None else None
    n = lib().orc_dedupe_keep_first(s, d, ww.ctypes.data if ww is not None else None, s.size)
    return s[:n].copy(), d[:n].copy(), (ww[:n].copy() if ww is not None else None)


def rmat_graph(scale, edge_factor, seed, weights=None, wseed=None):
    """Generated, permuted, deduplicated R-MAT edge list and its N (max id+1)."""
    s, d = rmat(scale, edge_factor, seed)
    s, d, _ = dedupe_keep_first(s, d)
    N = int(max(s.max(), d.max())) + 1
    return s, d, N


# --------------------------------------------------------------------------- closed forms

def dense_ref_fixed_point(N, outdeg, row_ptr, col, damp=0.85):
    """REF-mode fixed point: r = (1-d) (I - d A^T D^+)^-1 1, returned /N."""
    M = np.zeros((N, N))
    for v in range(N):
        for j in range(int(row_ptr[v]), int(row_ptr[v + 1])):
            u = int(col[j])
            M[v, u] += 1.0 / outdeg[u]
    r = np.linalg.solve(np.eye(N) - damp * M, (1 - damp) * np.ones(N))
    return r / N


def dense_stochastic(N, outdeg, row_ptr, col, damp=0.85, teleport=None, weights=None):
    """Stochastic PageRank with uniform dangling redistribution, dense solve."""
    t = np.full(N, 1.0 / N) if teleport is None else np.asarray(teleport, np.float64)
    M = np.zeros((N, N))
    for v in range(N):
        for j in range(int(row_ptr[v]), int(row_ptr[v + 1])):
            u = int(col[j])
            M[v, u] += (weights[j] if weights is not None else 1.0 / outdeg[u])
    dang = np.array([1.0 if outdeg[u] == 0 else 0.0 for u in range(N)])
    A = damp * (M + np.outer(np.ones(N), dang) / N)
    return np.linalg.solve(np.eye(N) - A, (1 - damp) * t)
