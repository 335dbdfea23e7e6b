import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIXTURES = ["v500_e500", "v1024_e4096", "v8192_e32768", "v1000_e10000", "v529_e3500"]
RMAT_SMALL = [8, 10, 12]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def load_golden(name: str):
    path = os.path.join(GOLDEN, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = load_golden(name)
        return cache[name]

    return get


@pytest.fixture(scope="session")
def gpu_ctx():
    import graphblox_amd as gbx

    print(f"[gpu] creating context on device 0 ({gbx.LIB_HIP})", flush=True)
    ctx = gbx.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture
def tuning(gpu_ctx):
    """Sets tuning fields on the shared context for one test (the context
    reads GBX_* environment overrides only when it is created), restoring
    the context's tuning afterwards."""
    saved = gpu_ctx.tuning()

    def set_fields(**fields):
        t = gpu_ctx.tuning()
        for k, v in fields.items():
            setattr(t, k, v)
        gpu_ctx.set_tuning(t)

    yield set_fields
    gpu_ctx.set_tuning(saved)


def l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).sum())
