import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "hpg-libs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_small():
    return load_golden("golden_small.npz")


@pytest.fixture(scope="session")
def golden_repeats():
    return load_golden("golden_repeats.npz")


def split_golden_hits(g, prefix="s1"):
    """Per-(read, strand) record lists of a golden set (prefix s1: BWSearch1CPU,
    s1v: BWSearch1VectorHelper on whole reads, s1w: on the s1w_bounds windows)."""
    counts = g[f"{prefix}_counts"]
    flat = g[f"{prefix}_hits"]
    out = {}
    o = 0
    for i in range(counts.shape[0]):
        for s in (0, 1):
            c = int(counts[i, s])
            out[(i, s)] = flat[o:o + c]
            o += c
    return out


@pytest.fixture(scope="session")
def gpu_engine():
    """The HIP engine on device 0.  GPU tests fail (not skip) without a GPU:
    they only run when selected with -m gpu on a GPU machine."""
    import hpgfm
    if hpgfm.device_count() < 1:
        pytest.fail("no GPU visible to the HIP engine")
    return hpgfm
