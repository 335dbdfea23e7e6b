"""Benchmark: pull PageRank on R-MAT scale 24 on one MI355X, and the other
workloads of the path beside it.

A step is one PageRank iteration (reference semantics, pageRank.c) over the
whole device-resident graph.  The headline runs the library's default method
for this size, the binned (propagation-blocking) iteration -- a scatter kernel
writes every edge's source contribution into destination-range bins and an
accumulate kernel sums the bins in LDS and applies the update -- with its
layout cut chosen by the library's autotuner on this graph; the timed
iterations are replayed as one captured HIP graph (`--eager`: launched one
by one; the other is reported among the forms).

Prints one JSON line: GTEPS (E edges per iteration / wall time per
iteration), the HBM roofline of the iteration (algorithmic bytes of the pull
formulation, 8E + 16N + 4, over the device time per iteration between the HIP
events that bracket the timed iterations on the library's stream), the error
of its ranks against the FP64 oracle, the oracle's OpenMP port of
pageRankPullGraphCSR timed on this host's cores on the same graph, wall time
to convergence, and `forms`: the pull kernel, BF16, the 32-vector batched
form, weighted PageRank and weighted SpMV on R-MAT s22, stochastic PageRank
to L1 < 1e-6 on s22, s16 beside the CPU port, the batch of 256 graphs, the
adjoint on s20 -- each with its error against FP64 (or the port) -- and the
edge-file ingest (read + build) at s22 beside the CPU path and at s24.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--scale 24]
"""
from __future__ import annotations

import argparse
import ctypes
import functools
import json
import os
import sys
import tempfile
import time

# the CPU baseline's OpenMP threads spread over the cores (read when the
# oracle library's OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "spread")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 24


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(workload_key: str):
    """HBM bytes per launch measured with rocprofv3 --pmc, if recorded."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("hbm_bytes_per_launch")


@functools.lru_cache(maxsize=None)
def host_cores():
    """Logical CPUs of the machine, CPUs this process may run on, physical cores.
    Counted once, at start-up: once an OpenMP region has run with
    OMP_PROC_BIND set, the runtime has bound this thread to one place."""
    logical = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = logical
    phys = set()
    try:
        pid = cid = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":")[1].strip()
                elif not line.strip() and pid is not None:
                    phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    return logical, allowed, len(phys) or logical


def baseline_threads():
    """The OpenMP threads of the CPU baseline: OMP_NUM_THREADS when set (the
    host's CPU share for this GPU), else every CPU this process may use."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.strip().isdigit() and int(env) > 0:
        return int(env)
    return host_cores()[1]


def csr_for_oracle(g):
    from oracle import oracle as O

    fwd = g.export(inverse=False)
    inv = g.export(inverse=True)
    mk = lambda c: O.OracleCSR(g.N, g.E, c["out_degree"], c["in_degree"], c["edges_idx"], c["src"], c["dst"], c["w"])
    return mk(fwd), mk(inv)


def fp64_errors(ranks, ref64):
    """L1 and largest relative error of ranks against an FP64 result."""
    r = np.asarray(ranks, np.float64)
    d = np.abs(r - ref64)
    den = np.maximum(np.abs(ref64), 1e-300)
    return {"l1_vs_fp64": float(d.sum()), "max_rel_vs_fp64": float((d / den).max())}


def cpu_baseline(fo, io, iterations: int, threads: int, gpu_ranks=None):
    """The oracle's restatement of pageRankPullGraphCSR (pageRank.c:245-368;
    its ranks, error and stopping iteration are bit-exact with the reference's
    own build, tests/test_oracle.py) on this graph's CSR arrays: `iterations`
    iterations with `threads` OpenMP threads, each iteration timed; the value
    is E / (median time of iterations 2..n).  With gpu_ranks (the measured
    GPU path's ranks after the same iterations) the record holds their L1
    distance from the port's ranks."""
    from oracle import oracle as O

    N, E = fo.N, fo.E
    logical, allowed, physical = host_cores()
    t0 = time.perf_counter()
    out = O.pr_pull_f32(fo, io, iterations, 0.0, nthreads=threads, timings=True)
    wall = time.perf_counter() - t0
    ts = np.asarray(out["iter_seconds"], np.float64)
    med = float(np.median(ts[1:])) if ts.size > 1 else float(ts[0])
    rec = dict(value=round(E / med / 1e9, 4), unit="GTEPS", cores=threads, kind="port",
               sample=f"{iterations} iterations of pull PageRank (FP32, reference semantics, the oracle's "
                      f"OpenMP restatement of pageRankPullGraphCSR) on the same R-MAT graph (N={N}, E={E}), "
                      f"{threads} OpenMP threads, OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', '')}; "
                      f"value = E / median of iterations 2..{iterations}",
               median_iter_ms=round(med * 1e3, 3), first_iter_ms=round(float(ts[0]) * 1e3, 3),
               wall_s=round(wall, 3), logical_cpus=logical, allowed_cpus=allowed, physical_cores=physical)
    if gpu_ranks is not None:
        rec["l1_gpu_vs_port"] = float(np.abs(np.asarray(gpu_ranks, np.float64) - out["ranks"]).sum())
    return rec, out["ranks"]


def timed_steps(pr, steps, graph_replay=False):
    """The timed region, after a warm-up: reset the session (its start event
    is recorded at the next launch), launch `steps` iterations back to back
    as a user runs them (or as one captured HIP graph).  Returns (device ms
    per iteration between the HIP events the library records on its stream
    before the first and after the last launch, host wall ms per iteration,
    ranks, report)."""
    ctx = pr.ctx
    pr.reset()
    ctx.sync()
    t0 = time.perf_counter()
    if graph_replay:
        pr.launch_graph(steps)
    else:
        pr.launch(steps)
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    ranks, rep = pr.read()
    rep0 = rep[0] if isinstance(rep, list) else rep
    return rep0["time_total_s"] * 1e3 / steps, wall_ms, ranks, rep


# --no-graphs: every launch one by one (for profiler counter runs, which
# have crashed inside a HIP graph launch)
NO_GRAPHS = False


def launch_iters(pr, n):
    if NO_GRAPHS:
        pr.launch(n)
    else:
        pr.launch_graph(n)


def time_session(pr, warmup, steps, graph_replay=False):
    graph_replay = graph_replay and not NO_GRAPHS
    pr.reset()
    if graph_replay:
        pr.launch_graph(steps)  # captured here, replayed in the timed region
    else:
        pr.launch(warmup)
    return timed_steps(pr, steps, graph_replay)


def time_form(g, warmup, steps, graph_replay=True, **kw):
    """Warm up, then time `steps` iterations (timed_steps), replayed as a
    captured HIP graph like the headline (launched one by one with
    --no-graphs or graph_replay=False).  Returns (kernel_ms, wall_ms, ranks,
    session)."""
    import graphblox_amd as gbx

    k = kw.pop("num_vectors", 0)
    seeds = kw.pop("seed_sets", None)
    if k:
        pr = g.ppr_session(k, seeds, iterations=warmup + steps, epsilon=0.0, stop=gbx.STOP_NONE)
    else:
        pr = g.session(iterations=warmup + steps, epsilon=0.0, stop=gbx.STOP_NONE, **kw)
    kernel_ms, wall_ms, ranks, _ = time_session(pr, warmup, steps, graph_replay)
    return kernel_ms, wall_ms, ranks, pr


def form_record(name, E, N, k, kernel_ms, wall_ms, bytes_per_iter, base_gteps=None, **extra):
    gteps = k * E / (wall_ms * 1e-3) / 1e9
    rec = {"form": name, "num_vectors": k, "num_vertices": N, "num_edges": E,
           "ms_per_iter": round(wall_ms, 5), "kernel_ms": round(kernel_ms, 5),
           "gteps": round(gteps, 2), "algorithmic_bytes_per_iter": bytes_per_iter,
           "hbm_frac": round(bytes_per_iter / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if base_gteps:
        rec["speedup_vs_fp32_same_graph"] = round(gteps / base_gteps, 3)
    rec.update(extra)
    log(f"{name}: {rec}")
    return rec


def calibrate_copy(gib: float):
    """A plain streaming copy of known size: y = x * 1 over gib GiB of float32
    (reads gib GiB, writes gib GiB), three launches."""
    import torch

    n = int(gib * (1 << 30)) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": n * 4, "calibrate_copy_bytes_written": n * 4}))


# ---------------------------------------------------------------- forms


def forms_s24(gbx, O, g, r64_ref, args, headline_gteps, method):
    """The pull kernel (or the binned iteration) and BF16 with both methods on
    the headline graph, after the same iterations as the headline."""
    N, E = g.N, g.E
    out = {}
    other = "pull" if method == "binned" else "binned"
    other_m = gbx.METHOD_PULL if other == "pull" else gbx.METHOD_BINNED
    km, wm, r1, s1 = time_form(g, args.warmup, args.steps, mode=gbx.PR_REF, method=other_m)
    s1.free()
    base_pull = E / (wm * 1e-3) / 1e9 if other == "pull" else headline_gteps
    out[f"{other}_f32"] = form_record(f"FP32 {other} iteration, same graph", E, N, 1, km, wm, 8 * E + 16 * N + 4,
                                      **fp64_errors(r1, r64_ref))
    for mname, m in (("binned", gbx.METHOD_BINNED), ("pull", gbx.METHOD_PULL)):
        km, wm, r16, s16 = time_form(g, args.warmup, args.steps, mode=gbx.PR_REF, precision=gbx.BF16, method=m)
        s16.free()
        rec = form_record(f"BF16 contributions, {mname} iteration, same graph", E, N, 1, km, wm,
                          6 * E + 18 * N + 4, headline_gteps if mname == "binned" else base_pull,
                          **fp64_errors(r16, r64_ref))
        rec["speedup_vs_plain_pull_fp32"] = round(rec["gteps"] / base_pull, 3)
        out[f"bf16_{mname}"] = rec
    return out


def forms_s22(gbx, O, ctx, args):
    """R-MAT s22: plain stochastic PageRank (pull, binned), the 32-vector
    batched form, stochastic PageRank to L1 < 1e-6, and uniform(0,1)-weighted
    PageRank and SpMV."""
    out = {}
    gb = ctx.rmat_graph(min(args.scale, 22), args.edge_factor, SEED)
    Nb, Eb = gb.N, gb.E
    fo, io = csr_for_oracle(gb)
    rp = O.row_ptr_from(io)
    r64s, _, _ = O.pr_f64(Nb, fo.out_degree, rp, io.dst, mode=1, iterations=args.steps)
    km, wm, rr, s = time_form(gb, args.warmup, args.steps, mode=gbx.PR_STOCHASTIC, method=gbx.METHOD_PULL)
    s.free()
    base32 = Eb / (wm * 1e-3) / 1e9
    out["fp32_pull_same_graph_as_batched"] = form_record("plain FP32 pull (stochastic)", Eb, Nb, 1, km, wm,
                                                         8 * Eb + 16 * Nb + 4, **fp64_errors(rr, r64s))
    km, wm, rr, s = time_form(gb, args.warmup, args.steps, mode=gbx.PR_STOCHASTIC, method=gbx.METHOD_BINNED)
    s.free()
    out["fp32_binned_same_graph_as_batched"] = form_record("FP32 binned (stochastic)", Eb, Nb, 1, km, wm,
                                                           8 * Eb + 16 * Nb + 4, **fp64_errors(rr, r64s))
    rng = np.random.default_rng(SEED)
    sets = [rng.choice(Nb, int(rng.integers(1, 1001)), replace=False) for _ in range(32)]
    # the batched form's lanes per row, hub cut and blocks per CU, searched
    # by the library's autotuner on this graph with 32 vectors (--no-autotune:
    # the library's defaults)
    if args.no_autotune:
        best = ctx.tuning()
    else:
        t0 = time.perf_counter()
        best, _ = ctx.autotune(gb, 32, 5, iterations=10, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE)
        log(f"batched form autotuned in {time.perf_counter() - t0:.2f} s")
    ppr_tuned = {k: v for k, v in best.as_dict().items() if k.startswith("ppr")}
    log(f"batched form tuning: {ppr_tuned}")
    kb, wb, rb, sb = time_form(gb, args.warmup, args.steps, num_vectors=32, seed_sets=sets)
    sb.free()
    errs = []
    for q in range(0, 32, 8):
        t = np.zeros(Nb)
        t[sets[q]] = 1.0 / len(sets[q])
        r64q, _, _ = O.pr_f64(Nb, fo.out_degree, rp, io.dst, mode=1, iterations=args.steps, teleport=t)
        errs.append(fp64_errors(rb[:, q], r64q))
    out["batched_ppr32"] = form_record(
        "32 personalised vectors, one multi-vector launch per iteration, the iterations replayed as a HIP graph",
        Eb, Nb, 32, kb, wb,
        4 * (Nb + 1) + 4 * Nb + 4 * Eb + 32 * (4 * Eb + 8 * Nb), base32,
        l1_vs_fp64=max(e["l1_vs_fp64"] for e in errs), max_rel_vs_fp64=max(e["max_rel_vs_fp64"] for e in errs),
        fp64_checked_vectors="4 of 32 (0, 8, 16, 24)", autotuned=ppr_tuned)
    out["batched_ppr32"]["speedup_vs_binned_fp32_same_graph"] = round(
        out["batched_ppr32"]["gteps"] / out["fp32_binned_same_graph_as_batched"]["gteps"], 3)
    # stochastic PageRank to L1 < 1e-6 (at most 100 iterations), wall time from reset to read-back
    conv = gb.session(iterations=100, epsilon=1e-6, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
    ctx.sync()
    t0 = time.perf_counter()
    conv.reset()
    conv.launch(100)
    rc, repc = conv.read()
    wall = (time.perf_counter() - t0) * 1e3
    conv.free()
    r64c, it64, _ = O.pr_f64(Nb, fo.out_degree, rp, io.dst, mode=1, iterations=100, epsilon=1e-6)
    out["stochastic_to_l1_1e-6"] = {"form": "stochastic PageRank to L1 < 1e-6, binned", "num_vertices": Nb,
                                    "num_edges": Eb, "iterations": repc["executed"], "fp64_iterations": it64,
                                    "wall_ms": round(wall, 3), "rank_sum": repc["rank_sum"], **fp64_errors(rc, r64c)}
    log(f"s22 to L1<1e-6: {out['stochastic_to_l1_1e-6']}")
    gb.free()
    # uniform(0,1) weights: weighted stochastic PageRank and weighted SpMV, both methods
    gw = ctx.rmat_graph(min(args.scale, 22), args.edge_factor, SEED, weight_kind=1, weight_seed=SEED,
                        flags=gbx.GRAPH_WEIGHTED)
    Nw, Ew = gw.N, gw.E
    fo, io = csr_for_oracle(gw)
    rp = O.row_ptr_from(io)
    S = np.bincount(fo.src, weights=fo.w.astype(np.float64), minlength=Nw)
    wn = np.where(S[io.dst] > 0, io.w.astype(np.float64) / np.maximum(S[io.dst], 1e-300), 0.0)
    r64w, _, _ = O.pr_f64(Nw, fo.out_degree, rp, io.dst, mode=1, iterations=args.steps, weights=wn)
    for mname, m in (("binned", gbx.METHOD_BINNED), ("pull", gbx.METHOD_PULL)):
        km, wm, rw, s = time_form(gw, args.warmup, args.steps, mode=gbx.PR_STOCHASTIC, weighted=True, method=m)
        s.free()
        out[f"weighted_pr_{mname}"] = form_record(f"weighted stochastic PageRank, uniform(0,1) weights, {mname}",
                                                  Ew, Nw, 1, km, wm, 12 * Ew + 16 * Nw + 4, **fp64_errors(rw, r64w))
    xo, yo = O.spmv_pull_f32(fo, io, iterations=1, weights=io.w, nthreads=0)
    tune0 = ctx.tuning()
    for mname, thr in (("binned", 0), ("pull", 1 << 30)):
        t = ctx.tuning()
        t.binned_min_vertices = thr  # gbx_spmv_pull takes the binned iteration from this many vertices
        ctx.set_tuning(t)
        gw.spmv(iterations=args.warmup, weighted=True)
        _, y, sec = gw.spmv(iterations=args.steps, weighted=True)
        ms = sec * 1e3 / args.steps
        _, y1, _ = gw.spmv(iterations=1, weighted=True)
        rel = float(np.abs(y1.astype(np.float64) - yo).sum() / np.abs(yo.astype(np.float64)).sum())
        out[f"weighted_spmv_{mname}"] = form_record(f"weighted SpMV y = A^T x, uniform(0,1) weights, {mname}",
                                                    Ew, Nw, 1, ms, ms, 12 * Ew + 8 * Nw + 4,
                                                    rel_l1_vs_port=rel, timing="device events around the passes")
    ctx.set_tuning(tune0)
    gw.free()
    return out


def forms_s16(gbx, O, ctx, args):
    """R-MAT s16 (65,536 vertices): the GPU beside the CPU port, 20 iterations."""
    g = ctx.rmat_graph(16, args.edge_factor, SEED)
    N, E = g.N, g.E
    fo, io = csr_for_oracle(g)
    km, wm, r, s = time_form(g, args.warmup, args.steps, mode=gbx.PR_REF)
    s.free()
    rec_cpu, port = cpu_baseline(fo, io, args.steps, baseline_threads(), r)
    r64, _, _ = O.pr_f64(N, fo.out_degree, O.row_ptr_from(io), io.dst, mode=0, iterations=args.steps)
    rec = form_record("R-MAT s16, reference semantics, GPU (default method) beside the CPU port", E, N, 1, km, wm,
                      8 * E + 16 * N + 4, **fp64_errors(r, r64))
    rec["cpu_port"] = rec_cpu
    rec["gpu_vs_cpu_port_speedup"] = round(rec["gteps"] / rec_cpu["value"], 2)
    ks, ws, rs, ss = time_form(g, args.warmup, args.steps, mode=gbx.PR_STOCHASTIC)
    ss.free()
    r64s, _, _ = O.pr_f64(N, fo.out_degree, O.row_ptr_from(io), io.dst, mode=1, iterations=args.steps)
    rec["stochastic"] = form_record("R-MAT s16, stochastic (uniform dangling mass)", E, N, 1, ks, ws,
                                    8 * E + 16 * N + 4, **fp64_errors(rs, r64s))
    g.free()
    return rec


def forms_batch(gbx, O, ctx, args):
    """256 independent R-MAT graphs, 2^10..2^18 vertices (log-uniform), edge
    factor 8..32: every graph's stochastic PageRank in the same kernels each
    iteration (the binned iteration at this size: scatter, accumulate with
    each unit's graph coefficients, per-graph reduction), the iterations
    replayed as one captured HIP graph, the layout cut autotuned on the
    batch (--no-autotune: the library's default cut); the pull tiles beside
    it."""
    rng = np.random.default_rng(256)
    scales = rng.integers(10, 19, 256)
    efs = rng.integers(8, 33, 256)
    members = [ctx.rmat_graph(int(s), int(f), 5000 + k) for k, (s, f) in enumerate(zip(scales, efs))]
    b = ctx.graph_batch(members)
    off = b.batch_offsets()
    N, E = b.N, b.E
    tune0 = ctx.tuning()
    if args.no_autotune:  # the library's default cut (counter runs want one layout)
        best, tuned_ms = tune0, None
    else:
        best, tuned_ms = ctx.autotune(b, 0, 5, iterations=10, epsilon=0.0, mode=gbx.PR_STOCHASTIC,
                                      stop=gbx.STOP_NONE)
    s = b.session(iterations=args.steps, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE)
    km, wm, ranks, reps = time_session(s, args.warmup, args.steps, graph_replay=True)
    s.free()
    errs = []
    for k in (0, 85, 170, 255):
        m = members[k]
        fo, io = csr_for_oracle(m)
        r64, _, _ = O.pr_f64(m.N, fo.out_degree, O.row_ptr_from(io), io.dst, mode=1, iterations=args.steps)
        errs.append(fp64_errors(ranks[off[k]:off[k + 1]], r64))
    # to convergence (each graph its own L1 < 1e-7 stop), one replayed graph of 100 iterations
    c = b.session(iterations=100, epsilon=1e-7, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_L1)
    c.reset()
    launch_iters(c, 100)
    c.read()
    ctx.sync()
    t0 = time.perf_counter()
    c.reset()
    launch_iters(c, 100)
    _, creps = c.read()
    conv_ms = (time.perf_counter() - t0) * 1e3
    c.free()
    rec = form_record("256 R-MAT graphs (2^10..2^18 vertices, edge factor 8..32), all graphs in each iteration's "
                      "kernels (the library's default method for this size), replayed as a HIP graph", E, N, 1, km,
                      wm, 8 * E + 16 * N + 4 * 256,
                      l1_vs_fp64=max(e["l1_vs_fp64"] for e in errs),
                      max_rel_vs_fp64=max(e["max_rel_vs_fp64"] for e in errs),
                      fp64_checked_graphs="4 of 256 (0, 85, 170, 255)",
                      autotuned={**{k: v for k, v in best.as_dict().items() if k.startswith("binned")},
                                 "trial_ms": None if tuned_ms is None else round(tuned_ms, 5)},
                      convergence={"epsilon": 1e-7, "wall_ms": round(conv_ms, 3),
                                   "iterations_max": max(r["executed"] for r in creps),
                                   "all_converged": all(r["converged"] for r in creps)})
    ctx.set_tuning(tune0)
    sp = b.session(iterations=args.steps, epsilon=0.0, mode=gbx.PR_STOCHASTIC, stop=gbx.STOP_NONE,
                   method=gbx.METHOD_PULL)
    kp, wp, rp_, _ = time_session(sp, args.warmup, args.steps, graph_replay=True)
    sp.free()
    rec["pull_tiles"] = {"ms_per_iter": round(wp, 5), "kernel_ms": round(kp, 5),
                         "gteps": round(E / (wp * 1e-3) / 1e9, 2),
                         "hbm_frac": round((8 * E + 16 * N + 4 * 256) / (kp * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "l1_vs_default_method": float(np.abs(rp_.astype(np.float64) - ranks).sum() / 256)}
    b.free()
    for m in members:
        m.free()
    return rec


class _EdgeList(ctypes.Structure):
    """struct EdgeList (edgeList.h:12-25) as the host adapter returns it."""
    _fields_ = [("num_edges", ctypes.c_uint32), ("num_vertices", ctypes.c_uint32), ("avg_degree", ctypes.c_uint32),
                ("max_weight", ctypes.c_float), ("edges_array_weight", ctypes.c_void_p),
                ("edges_array_src", ctypes.c_void_p), ("edges_array_dest", ctypes.c_void_p),
                ("mask_array", ctypes.c_void_p), ("label_array", ctypes.c_void_p),
                ("inverse_label_array", ctypes.c_void_p)]


def cpu_ingest_baseline(O, path, N_gpu, E_gpu):
    """The CPU side of the edge-file ingest: the host adapter's
    readEdgeListsbin (the reference's serial record loop, edgeList.c:551-677)
    on the same file, then the oracle's restatement of the DIRECTED=1 build
    (count sort by source and destination, countsort.c:262-270, the vertex
    map, the transpose and its sort; serial C) on the arrays it read."""
    L = ctypes.CDLL(os.path.join(ROOT, "graphblox_amd", "libgraphblox_hip.so"))
    L.readEdgeListsbin.restype = ctypes.POINTER(_EdgeList)
    L.readEdgeListsbin.argtypes = [ctypes.c_char_p, ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint32]
    L.freeEdgeList.argtypes = [ctypes.POINTER(_EdgeList)]
    t0 = time.perf_counter()
    ep = L.readEdgeListsbin(os.fsencode(path), 0, 0, 0)
    t_read = time.perf_counter() - t0
    e = ep.contents
    E, N = e.num_edges, e.num_vertices
    s = np.ctypeslib.as_array(ctypes.cast(e.edges_array_src, ctypes.POINTER(ctypes.c_uint32)), (E,)).copy()
    d = np.ctypeslib.as_array(ctypes.cast(e.edges_array_dest, ctypes.POINTER(ctypes.c_uint32)), (E,)).copy()
    L.freeEdgeList(ep)
    t0 = time.perf_counter()
    O.build_graph(s, d, None, N, mode=1)
    t_build = time.perf_counter() - t0
    assert (N, E) == (N_gpu, E_gpu)
    return {"read_s": round(t_read, 3), "build_s": round(t_build, 3), "total_s": round(t_read + t_build, 3),
            "threads": 1, "kind": "port",
            "sample": "the same file: the host adapter's readEdgeListsbin, then the oracle's serial restatement "
                      "of the DIRECTED=1 CSR + transpose build, sort mode 1"}


def forms_ingest(gbx, O, ctx, args, tmpdir):
    """An R-MAT edge file (.bin records, written to local storage and read
    back from the page cache) through the GPU ingest: gbx_edges_read_file
    (the records to the device through pinned 64 MiB pieces, deinterleave,
    max id) and gbx_graph_create_from_edges (sort by source and
    destination, the vertex map, the transpose); wall time of each, at s22
    beside the CPU path on the same file and at s24."""
    out = {}
    for scale in sorted({min(22, args.scale), args.scale}):
        s, d, _ = ctx.rmat_edges(scale, args.edge_factor, SEED)
        path = os.path.join(tmpdir, f"rmat_s{scale}.bin")
        np.ascontiguousarray(np.stack([s, d], axis=1)).tofile(path)
        del s, d
        nbytes = os.path.getsize(path)
        ctx.sync()
        t0 = time.perf_counter()
        ed = ctx.edges_read_file(path)
        t_read = time.perf_counter() - t0
        N, E = ed.N, ed.E
        t0 = time.perf_counter()
        g = ed.graph(sort_mode=gbx.SORT_RADIX_SRC_DST)
        ctx.sync()
        t_build = time.perf_counter() - t0
        ed.free()
        g.free()
        rec = {"file_bytes": nbytes, "num_vertices": N, "num_edges": E, "read_s": round(t_read, 4),
               "read_gbs": round(nbytes / t_read / 1e9, 2), "build_s": round(t_build, 4),
               "total_s": round(t_read + t_build, 4)}
        if scale == min(22, args.scale) and not args.no_cpu:
            rec["cpu"] = cpu_ingest_baseline(O, path, N, E)
            rec["gpu_vs_cpu_speedup"] = round(rec["cpu"]["total_s"] / rec["total_s"], 1)
        os.remove(path)
        out[f"rmat_s{scale}"] = rec
        log(f"ingest s{scale}: {rec}")
    return out


def forms_adjoint(gbx, O, ctx, args):
    """Weighted PageRank on R-MAT s20 with lognormal(0,1) weights and its
    adjoint (loss weights N(0,1)): forward and adjoint fixed points to
    1e-7, then dL/dW; ms per iteration of each and wall time to the gradient,
    both methods; the gradient against the FP64 adjoint."""
    g = ctx.rmat_graph(20, args.edge_factor, SEED, weight_kind=2, weight_seed=SEED, flags=gbx.GRAPH_WEIGHTED)
    N, E = g.N, g.E
    gl = np.random.default_rng(SEED).standard_normal(N).astype(np.float32)
    inv = g.export(inverse=True)
    rp = np.zeros(N + 1, np.uint64)
    np.cumsum(inv["out_degree"], out=rp[1:])
    p, lam, gw64, _ = O.adjoint_f64(N, rp, inv["dst"], inv["w"].astype(np.float64), gl, iterations=2000, tol=1e-14)
    f = g.export()
    # the forward-CSR position of every inverse-CSR edge (u -> v is (src u, dst v) forward, (v, u) inverse)
    key_fwd = f["src"].astype(np.uint64) * np.uint64(N) + f["dst"].astype(np.uint64)
    key = inv["dst"].astype(np.uint64) * np.uint64(N) + inv["src"].astype(np.uint64)
    order = np.argsort(key_fwd)
    perm = order[np.searchsorted(key_fwd[order], key)]
    assert np.array_equal(key_fwd[perm], key)
    del key_fwd, key, order
    out = {}
    for mname, m in (("binned", gbx.METHOD_BINNED), ("pull", gbx.METHOD_PULL)):
        g.adjoint(gl, iterations=300, epsilon=1e-7, method=m)  # plans, layouts and weights built
        t0 = time.perf_counter()
        o = g.adjoint(gl, iterations=300, epsilon=1e-7, method=m)
        wall = (time.perf_counter() - t0) * 1e3
        fr, ar = o["fwd_report"], o["adj_report"]
        gw = o["grad_w"][perm].astype(np.float64)
        out[mname] = {"forward_iterations": fr["executed"], "adjoint_iterations": ar["executed"],
                      "ms_per_forward_iteration": round(fr["time_total_s"] * 1e3 / max(fr["executed"], 1), 5),
                      "ms_per_adjoint_iteration": round(ar["time_total_s"] * 1e3 / max(ar["executed"], 1), 5),
                      "wall_ms_to_gradient": round(wall, 3),
                      "grad_w_rel_l2_vs_fp64": float(np.linalg.norm(gw - gw64) / np.linalg.norm(gw64)),
                      "lambda_rel_l1_vs_fp64": float(np.abs(o["lam"] - lam).sum() / np.abs(lam).sum())}
        log(f"adjoint s20 {mname}: {out[mname]}")
    g.free()
    return {"form": "adjoint of weighted stochastic PageRank, R-MAT s20, lognormal weights", "num_vertices": N,
            "num_edges": E, "algorithmic_bytes_per_iteration": 12 * E + 16 * N + 4, **out}


def convergence(gbx, ctx, args, g, method, m):
    """Reference semantics, stop when no vertex changes by 1e-4 or more
    (pageRank.c:327-339), at most 100 iterations; wall time from reset to
    read-back of the ranks, on the graph whose plan and binned layout exist
    (warm: the 100 iterations replayed as a HIP graph, like the headline),
    and one-shot on a graph just built, which pays both and launches its
    iterations eagerly, 8 between stop checks (cold: gbx_pr_pull, the
    drop-in call)."""
    conv_s = g.session(iterations=100, epsilon=1e-4, mode=gbx.PR_REF, stop=gbx.STOP_ACTIVE, method=m)
    conv_s.reset()
    launch_iters(conv_s, 100)  # the 100 iterations captured as a HIP graph once, outside the timing
    conv_s.read()
    ctx.sync()
    t0 = time.perf_counter()
    conv_s.reset()
    launch_iters(conv_s, 100)  # iterations after the stop rule fired exit at once
    _, conv_rep = conv_s.read()
    warm_ms = (time.perf_counter() - t0) * 1e3
    conv_s.free()
    g2 = ctx.rmat_graph(args.scale, args.edge_factor, SEED + 1)
    ctx.sync()
    t0 = time.perf_counter()
    _, rep2 = g2.pagerank(iterations=100, epsilon=1e-4, method=m)  # gbx_pr_pull: plan, layout, iterations, read
    cold_ms = (time.perf_counter() - t0) * 1e3
    g2.free()
    rec = {"epsilon": 1e-4, "stop": "no vertex with |dr| >= epsilon", "iterations": conv_rep["iterations"],
           "executed": conv_rep["executed"], "converged": bool(conv_rep["converged"]),
           "wall_ms": round(warm_ms, 3),
           "wall_ms_one_shot_with_plan_and_layout_build": round(cold_ms, 3),
           "one_shot_graph": f"R-MAT s{args.scale} seed {SEED + 1} (same shape, a new graph object)",
           "one_shot_iterations": rep2["iterations"]}
    log(f"convergence ({method}): {rec}")
    return rec


def main():
    host_cores()  # before any OpenMP region binds this thread
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: OMP_NUM_THREADS, else every allowed CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="the headline only")
    ap.add_argument("--tmpdir", default=None, help="where the ingest measurement writes its edge files")
    ap.add_argument("--eager", action="store_true", help="headline iterations launched one by one, not replayed")
    ap.add_argument("--no-graphs", action="store_true",
                    help="no HIP graph launch anywhere (profiler counter runs); implies --eager")
    ap.add_argument("--no-autotune", action="store_true", help="the library's default layout cut")
    ap.add_argument("--method", choices=("binned", "pull"), default="binned",
                    help="iteration of the headline (binned: the library default at this size)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="run only a plain device copy of this many GiB (reference for PMC byte counters)")
    args = ap.parse_args()
    if args.no_graphs:
        global NO_GRAPHS
        NO_GRAPHS = True
        args.eager = True
    if args.calibrate_copy > 0:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU: sharding over several GPUs is "
                                   "designed, not built, so --gpus must be 1"}))
        sys.exit(2)

    import graphblox_amd as gbx
    from oracle import oracle as O

    ctx = gbx.Context(0)
    t0 = time.perf_counter()
    g = ctx.rmat_graph(args.scale, args.edge_factor, SEED)
    N, E = g.N, g.E
    log(f"R-MAT s{args.scale}: N={N} E={E} built on GPU in {time.perf_counter() - t0:.2f}s")
    method = gbx.METHOD_BINNED if args.method == "binned" else gbx.METHOD_PULL
    tuned = None
    tune_default = ctx.tuning()
    if not args.no_autotune:
        t0 = time.perf_counter()
        best, best_ms = ctx.autotune(g, 0, 5, iterations=10, epsilon=0.0, mode=gbx.PR_REF, stop=gbx.STOP_NONE,
                                     method=method)
        d = best.as_dict()
        tuned = {k: d[k] for k in d if (k.startswith("binned") if args.method == "binned" else
                                        not k.startswith(("binned", "ppr")))}
        tuned["trial_ms"] = round(best_ms, 5)
        tuned["search_s"] = round(time.perf_counter() - t0, 2)
        log(f"autotuned: {tuned}")
    t0 = time.perf_counter()
    pr = g.session(iterations=args.warmup + args.steps, epsilon=0.0, mode=gbx.PR_REF, stop=gbx.STOP_NONE,
                   method=method)
    log(f"{args.method} iteration set up in {time.perf_counter() - t0:.2f}s")
    # the timed iterations replayed as one captured HIP graph (captured and run once before)
    kernel_ms, ms_per_step, ranks, rep = time_session(pr, args.warmup, args.steps, graph_replay=not args.eager)
    assert rep["executed"] == args.steps
    assert np.isfinite(ranks).all()
    value = E / (ms_per_step * 1e-3) / 1e9
    bytes_per_launch = 8 * E + 16 * N + 4
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    wk = f"pr_{args.method}_f32_rmat_s{args.scale}"
    traffic = traffic_from_profiles(wk)
    log(f"{ms_per_step:.4f} ms/iter wall, kernels {kernel_ms:.4f} ms, {value:.1f} GTEPS, "
        f"{achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.3f} of peak)")
    fo, io = csr_for_oracle(g)
    r64, _, _ = O.pr_f64(N, fo.out_degree, O.row_ptr_from(io), io.dst, mode=0, iterations=args.steps)
    headline_err = fp64_errors(ranks, r64)
    log(f"headline vs FP64: {headline_err}")
    cpu = None
    if not args.no_cpu:
        log("CPU baseline ...")
        cpu, port = cpu_baseline(fo, io, args.cpu_iters, args.cpu_threads or baseline_threads(),
                                 ranks if args.cpu_iters == args.steps else None)
        log(f"CPU baseline {cpu['value']} GTEPS ({cpu['kind']}, {cpu['cores']} threads)")
    conv = convergence(gbx, ctx, args, g, args.method, method)
    forms = None
    if not args.no_forms:
        forms = forms_s24(gbx, O, g, r64, args, value, args.method)
        if not NO_GRAPHS:
            km, wm, re_, _ = time_session(pr, args.warmup, args.steps, graph_replay=args.eager)
            forms[f"{args.method}_f32_" + ("replayed" if args.eager else "eager")] = form_record(
                f"the headline's iteration {'replayed as a HIP graph' if args.eager else 'launched one by one'}",
                E, N, 1, km, wm, bytes_per_launch, **fp64_errors(re_, r64))
        del fo, io
        pr.free()
        g.free()
        ctx.set_tuning(tune_default)  # the cut tuned for s24 is not the other graphs' cut
        forms.update(forms_s22(gbx, O, ctx, args))
        forms["rmat_s16_gpu_and_cpu"] = forms_s16(gbx, O, ctx, args)
        forms["graph_batch_256"] = forms_batch(gbx, O, ctx, args)
        forms["adjoint_s20"] = forms_adjoint(gbx, O, ctx, args)
        with tempfile.TemporaryDirectory(dir=args.tmpdir) as td:
            forms["ingest"] = forms_ingest(gbx, O, ctx, args, td)
    out = {
        "metric": "GTEPS & % HBM roofline, pull PageRank R-MAT s24, 1 MI355X, vs plain FP32",
        "value": round(value, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT generated on the GPU)",
        "config": {
            "workload": f"pull PageRank, reference semantics (pageRank.c), FP32, R-MAT scale {args.scale} "
                        f"edge factor {args.edge_factor} a=.57 b=.19 c=.19 d=.05, ids permuted, "
                        f"self loops and duplicates removed; {args.method} iteration",
            "scale": args.scale, "edge_factor": args.edge_factor, "seed": SEED,
            "num_vertices": N, "num_edges": E, "parallelism": "single GPU", "method": args.method,
            "launch": "launched one by one" if args.eager else "the timed iterations replayed as one captured HIP graph",
            "autotuned": tuned,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": ("bin_scatter_runs_kernel<float> + bin_accum_kernel<K_PR_REF, float> (one iteration)"
                       if args.method == "binned" else "pr_iter_kernel<K_PR_REF>"),
            "kernel_ms": round(kernel_ms, 5),
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
        "error_vs_fp64": headline_err,
        "convergence": conv,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    if args.no_forms:
        pr.free()
        g.free()
    ctx.close()


if __name__ == "__main__":
    main()
