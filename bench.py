"""Benchmark: FP32 pull PageRank on R-MAT scale 24 on one MI355X.

A step is one PageRank iteration (reference semantics, pageRank.c) over the
whole device-resident graph.  The headline runs the library's default method
for this size, the binned (propagation-blocking) iteration: a scatter kernel
writes every edge's source contribution into destination-range bins and an
accumulate kernel sums the bins in LDS and applies the update, with the same
ranks as the pull iteration (within 1e-5 L1).  The plain pull kernel (one
gather per edge) is reported beside it, with the BF16 and 32-vector forms.

Prints one JSON line: GTEPS (E edges per iteration / wall time per
iteration), the HBM roofline of the iteration (algorithmic bytes of the pull
formulation, 8E + 16N + 4, over the device time per iteration between the
HIP events that bracket the timed iterations on the library's stream) and
the GraphBlox reference's OpenMP pull PageRank timed on this host's cores on
the same graph.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--scale 24]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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


def host_cores():
    """Logical CPUs of the machine, CPUs this process may run on, physical cores."""
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
    t0 = time.perf_counter()
    out = O.pr_pull_f32(fo, io, iterations, 0.0, nthreads=threads, timings=True)
    wall = time.perf_counter() - t0
    ts = np.asarray(out["iter_seconds"], np.float64)
    med = float(np.median(ts[1:])) if ts.size > 1 else float(ts[0])
    logical, allowed, physical = host_cores()
    rec = dict(value=round(E / med / 1e9, 4), unit="GTEPS", cores=threads, kind="port",
               sample=f"{iterations} iterations of pull PageRank (FP32, reference semantics, the oracle's "
                      f"OpenMP restatement of pageRankPullGraphCSR) on the same R-MAT graph (N={N}, E={E}), "
                      f"{threads} OpenMP threads, OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', '')}; "
                      f"value = E / median of iterations 2..{iterations}",
               median_iter_ms=round(med * 1e3, 3), first_iter_ms=round(float(ts[0]) * 1e3, 3),
               wall_s=round(wall, 3), logical_cpus=logical, allowed_cpus=allowed, physical_cores=physical)
    if gpu_ranks is not None:
        rec["l1_gpu_vs_port"] = float(np.abs(np.asarray(gpu_ranks, np.float64) - out["ranks"]).sum())
    return rec


def timed_steps(pr, steps):
    """The timed region, after a warm-up: reset the session (its start event
    is recorded at the next launch), launch `steps` iterations back to back
    as a user runs them.  Returns (device ms per iteration between the HIP
    events the library records on its stream before the first and after the
    last launch, host wall ms per iteration, ranks, report)."""
    ctx = pr.ctx
    pr.reset()
    ctx.sync()
    t0 = time.perf_counter()
    pr.launch(steps)
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    ranks, rep = pr.read()
    rep0 = rep[0] if isinstance(rep, list) else rep
    return rep0["time_total_s"] * 1e3 / steps, wall_ms, ranks, rep


def time_form(g, warmup, steps, **kw):
    """Warm up, then time `steps` iterations (timed_steps).  Returns
    (kernel_ms, wall_ms, session)."""
    k = kw.pop("num_vectors", 0)
    seeds = kw.pop("seed_sets", None)
    if k:
        pr = g.ppr_session(k, seeds, iterations=warmup + steps, epsilon=0.0, stop=2)
    else:
        pr = g.session(iterations=warmup + steps, epsilon=0.0, stop=2, **kw)
    pr.reset()
    pr.launch(warmup)
    kernel_ms, wall_ms, _, _ = timed_steps(pr, steps)
    return kernel_ms, wall_ms, pr


def form_record(name, E, N, k, kernel_ms, wall_ms, bytes_per_iter, base_gteps=None):
    gteps = k * E / (wall_ms * 1e-3) / 1e9
    rec = {"form": name, "num_vectors": k, "num_vertices": N, "num_edges": E,
           "ms_per_iter": round(wall_ms, 5), "kernel_ms": round(kernel_ms, 5),
           "gteps": round(gteps, 2), "algorithmic_bytes_per_iter": bytes_per_iter,
           "hbm_frac": round(bytes_per_iter / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if base_gteps:
        rec["speedup_vs_fp32_same_graph"] = round(gteps / base_gteps, 3)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: OMP_NUM_THREADS, else every allowed CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the BF16 and batched forms")
    ap.add_argument("--method", choices=("binned", "pull"), default="binned",
                    help="iteration of the headline (binned: the library default at this size)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="run only a plain device copy of this many GiB (reference for PMC byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU: sharding over several GPUs is "
                                   "designed, not built, so --gpus must be 1"}))
        sys.exit(2)

    import graphblox_amd as gbx

    ctx = gbx.Context(0)
    t0 = time.perf_counter()
    g = ctx.rmat_graph(args.scale, args.edge_factor, SEED)
    N, E = g.N, g.E
    log(f"R-MAT s{args.scale}: N={N} E={E} built on GPU in {time.perf_counter() - t0:.2f}s")
    method = gbx.METHOD_BINNED if args.method == "binned" else gbx.METHOD_PULL
    t0 = time.perf_counter()
    pr = g.session(iterations=args.warmup + args.steps, epsilon=0.0, mode=gbx.PR_REF, stop=gbx.STOP_NONE,
                   method=method)
    log(f"{args.method} iteration set up in {time.perf_counter() - t0:.2f}s")
    pr.reset()
    pr.launch(args.warmup)
    kernel_ms, ms_per_step, ranks, rep = timed_steps(pr, args.steps)
    assert rep["executed"] == args.steps
    assert np.isfinite(ranks).all()
    value = E / (ms_per_step * 1e-3) / 1e9
    bytes_per_launch = 8 * E + 16 * N + 4
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    wk = f"pr_{args.method}_f32_rmat_s{args.scale}"
    traffic = traffic_from_profiles(wk)
    log(f"{ms_per_step:.4f} ms/iter wall, kernels {kernel_ms:.4f} ms, {value:.1f} GTEPS, "
        f"{achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.3f} of peak)")
    forms = None
    if not args.no_forms:
        forms = {}
        other = "pull" if args.method == "binned" else "binned"
        other_m = gbx.METHOD_PULL if other == "pull" else gbx.METHOD_BINNED
        km, wm, s1 = time_form(g, args.warmup, args.steps, mode=gbx.PR_REF, method=other_m)
        r1, _ = s1.read()
        s1.free()
        base_pull = E / (wm * 1e-3) / 1e9 if other == "pull" else value
        forms[f"{other}_f32"] = form_record(f"FP32 {other} iteration, same graph", E, N, 1, km, wm,
                                            8 * E + 16 * N + 4)
        forms[f"{other}_f32"]["l1_vs_headline"] = float(np.abs(r1.astype(np.float64) - ranks).sum())
        log(f"{other} FP32 form: {forms[f'{other}_f32']}")
        for mname, m in (("binned", gbx.METHOD_BINNED), ("pull", gbx.METHOD_PULL)):
            km, wm, s16 = time_form(g, args.warmup, args.steps, mode=gbx.PR_REF, precision=gbx.BF16, method=m)
            s16.free()
            rec = form_record(f"BF16 contributions, {mname} iteration, same graph", E, N, 1, km, wm,
                              6 * E + 18 * N + 4, value)
            rec["speedup_vs_plain_pull_fp32"] = round(rec["gteps"] / base_pull, 3)
            forms[f"bf16_{mname}"] = rec
            log(f"BF16 {mname} form: {rec}")
        # 32 personalised vectors on R-MAT scale 22 (seed sets of 1..1000 vertices)
        gb = ctx.rmat_graph(min(args.scale, 22), args.edge_factor, SEED)
        Nb, Eb = gb.N, gb.E
        km32, wm32, s32 = time_form(gb, args.warmup, args.steps, mode=gbx.PR_STOCHASTIC, method=gbx.METHOD_PULL)
        s32.free()
        base32 = Eb / (wm32 * 1e-3) / 1e9
        forms["fp32_pull_same_graph_as_batched"] = form_record("plain FP32 pull (stochastic)", Eb, Nb, 1, km32,
                                                               wm32, 8 * Eb + 16 * Nb + 4)
        km32b, wm32b, s32b = time_form(gb, args.warmup, args.steps, mode=gbx.PR_STOCHASTIC,
                                       method=gbx.METHOD_BINNED)
        s32b.free()
        forms["fp32_binned_same_graph_as_batched"] = form_record("FP32 binned (stochastic)", Eb, Nb, 1, km32b,
                                                                 wm32b, 8 * Eb + 16 * Nb + 4)
        rng = np.random.default_rng(SEED)
        sets = [rng.choice(Nb, int(rng.integers(1, 1001)), replace=False) for _ in range(32)]
        kb, wb, sb = time_form(gb, args.warmup, args.steps, num_vectors=32, seed_sets=sets)
        rb, repb = sb.read()
        assert np.isfinite(rb).all()
        sb.free()
        forms["batched_ppr32"] = form_record("32 personalised vectors, one multi-vector launch", Eb, Nb, 32, kb, wb,
                                             4 * (Nb + 1) + 4 * Nb + 4 * Eb + 32 * (4 * Eb + 8 * Nb), base32)
        forms["batched_ppr32"]["speedup_vs_binned_fp32_same_graph"] = round(
            forms["batched_ppr32"]["gteps"] / forms["fp32_binned_same_graph_as_batched"]["gteps"], 3)
        log(f"batched form: {forms['batched_ppr32']}")
        gb.free()
    # wall time to convergence: reference semantics, stop when no vertex changes by 1e-4 or more
    # (pageRank.c:327-339), at most 100 iterations; reset, iterations and read-back timed
    conv_s = g.session(iterations=100, epsilon=1e-4, mode=gbx.PR_REF, stop=gbx.STOP_ACTIVE, method=method)
    ctx.sync()
    t0 = time.perf_counter()
    conv_s.reset()
    conv_s.launch(100)
    _, conv_rep = conv_s.read()
    conv_ms = (time.perf_counter() - t0) * 1e3
    conv_s.free()
    convergence = {"epsilon": 1e-4, "stop": "no vertex with |dr| >= epsilon", "iterations": conv_rep["iterations"],
                   "executed": conv_rep["executed"], "converged": bool(conv_rep["converged"]),
                   "wall_ms": round(conv_ms, 3)}
    log(f"convergence: {convergence}")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        log("CPU baseline ...")
        cs = g.session(iterations=args.cpu_iters, epsilon=0.0, mode=gbx.PR_REF, stop=gbx.STOP_NONE, method=method)
        cs.reset()
        cs.launch(args.cpu_iters)
        gpu_ranks, _ = cs.read()
        cs.free()
        fo, io = csr_for_oracle(g)
        threads = args.cpu_threads or baseline_threads()
        cpu = cpu_baseline(fo, io, args.cpu_iters, threads, gpu_ranks)
        log(f"CPU baseline {cpu['value']} GTEPS ({cpu['kind']}, {cpu['cores']} threads)")
        # the headline's ranks (after `steps` iterations) against the FP64 oracle
        r64, _, _ = O.pr_f64(N, fo.out_degree, O.row_ptr_from(io), io.dst, mode=0, iterations=args.steps)
        headline_err = fp64_errors(ranks, r64)
        log(f"headline vs FP64: {headline_err}")
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
        "error_vs_fp64": headline_err if cpu is not None else None,
        "convergence": convergence,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    pr.free()
    g.free()
    ctx.close()


if __name__ == "__main__":
    main()
