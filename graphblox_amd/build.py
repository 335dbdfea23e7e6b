"""Build recipe for the native libraries (run from any directory).

* ``graphblox_amd/libgbx_hip.so`` -- the HIP backend (csrc/*.hip) for gfx950,
  exporting the C ABI declared in ``include/gbx.h``;
* ``graphblox_amd/libgraphblox_hip.so`` -- the host adapter (host/*.c) that
  exposes the GraphBlox reference API (``include/graphblox_hip.h``) on top of
  the C ABI.

Both are built in-tree so the built files travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INC = os.path.join(ROOT, "include")

HIP_SOURCES = ["gbx_core.hip", "gbx_sort.hip", "gbx_rmat.hip", "gbx_pr.hip", "gbx_binned.hip", "gbx_ppr.hip",
               "gbx_batch.hip", "gbx_tune.hip", "gbx_ddpush.hip", "gbx_relabel.hip", "gbx_noalloc.hip",
               "gbx_ingest.hip"]
HOST_SOURCES = ["graphblox_hip.c", "graphstats_hip.c"]

LIB_HIP = os.path.join(PKG, "libgbx_hip.so")
LIB_HOST = os.path.join(PKG, "libgraphblox_hip.so")        # DIRECTED=1, WEIGHTED=0 layout
LIB_HOST_W = os.path.join(PKG, "libgraphblox_hip_w.so")    # DIRECTED=1, WEIGHTED=1 layout
# the adapter in each (DIRECTED, WEIGHTED) layout of the reference's compile-time switches
HOST_VARIANTS = {(1, 0): LIB_HOST, (1, 1): LIB_HOST_W,
                 (0, 0): os.path.join(PKG, "libgraphblox_hip_d0.so"),
                 (0, 1): os.path.join(PKG, "libgraphblox_hip_d0_w.so")}

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",  # keep the reference's separate multiply and add roundings
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd[:3])} ...")
    return res.stdout


def build_hip(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor

    hipcc = _hipcc()
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, src.replace(".hip", ".o")) for src in HIP_SOURCES]
    cmds = [[hipcc, *HIPCC_FLAGS, "-I", INC, "-c", os.path.join(CSRC, src), "-o", obj]
            for src, obj in zip(HIP_SOURCES, objs)]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", 0)) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for out in pool.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB_HIP + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB_HIP)
    return LIB_HIP


def build_host(verbose: bool = False) -> str:
    cc = shutil.which("gcc") or "cc"
    for (directed, weighted), lib in HOST_VARIANTS.items():
        tmp = lib + ".tmp"
        cmd = [cc, "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", "-Wextra", f"-DGBX_WEIGHTED={weighted}",
               f"-DGBX_DIRECTED={directed}", "-I", INC, *[os.path.join(HOST, s) for s in HOST_SOURCES],
               "-o", tmp, "-L", PKG, "-lgbx_hip", "-Wl,-rpath,$ORIGIN", "-lm"]
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        os.replace(tmp, lib)
    return LIB_HOST


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    if os.path.isdir(HOST) and all(os.path.exists(os.path.join(HOST, s)) for s in HOST_SOURCES):
        build_host(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB_HIP)
