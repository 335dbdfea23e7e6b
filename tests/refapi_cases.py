"""Cases shared by tests/golden/make_reference_api_goldens.py (which records
the reference's outputs for them) and the GPU tests that compare against
those outputs: the relabel orders and levels, the label file, and the seeded
rank-sort inputs."""
from __future__ import annotations

import hashlib

import numpy as np

# relabel orders of reorderGraphProcess (reorder.c:1556-1616) exercised per graph
LMODES = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14)
RELABEL_GRAPHS = ("rmat_s10", "fixture_v1024_e4096", "rmat_s12")
# composed levels (lmode, lmode_l2, lmode_l3) on rmat_s12, sort modes 0 and 1
LEVELS = ((4, 1, 8), (6, 0, 0), (0, 3, 0), (13, 9, 7))
# levels and -k run through the adapter's graphCSRPreProcessingStep on rmat_s10
API_LEVELS = (((4, 0, 0), 0), ((1, 8, 0), 1))


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def relabel_key(name, dflag, levels, sort, tag="") -> str:
    return f"relabel/{name}/d{dflag}/l{levels[0]}-{levels[1]}-{levels[2]}/s{sort}{tag}"


def file_labels(N: int) -> np.ndarray:
    """The label file of the order-11 case: a seeded permutation of [0, N)."""
    return np.random.default_rng(5).permutation(N).astype(np.uint32)


def rank_sort_cases():
    """(description, ranks, labels, inside the range C defines the float to
    u32 key conversion for), drawn in a fixed order from one seeded
    generator."""
    rng = np.random.default_rng(20261203)
    n = 200000
    specs = [
        ("normalised ranks of 1/N scale", lambda: (rng.pareto(1.2, n) / n).astype(np.float32), True),
        ("quantised values, many ties", lambda: (np.round(rng.random(n) * 50) / 64).astype(np.float32), True),
        ("up to 4095", lambda: (rng.random(n) * 4095).astype(np.float32), True),
        ("tiny", lambda: np.full(7, 1e-9, np.float32), True),
        ("beyond 2^12: keys wrap at 2^32", lambda: (rng.random(n) * 9000).astype(np.float32), False),
        ("some negative", lambda: (rng.random(n) - 0.1).astype(np.float32), False),
    ]
    for what, make, defined in specs:
        ranks = make()
        labels = rng.permutation(ranks.size).astype(np.uint32)
        yield what, ranks, labels, defined
