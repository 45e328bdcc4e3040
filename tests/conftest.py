import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def scanner():
    import epigpu_amd as E
    s = E.Scanner(0)
    yield s
    s.close()


def make_dataset(seed, nid, nsnp, maf=(0.05, 0.5), gmiss=0.0, dist=0, pmiss=0.0, sentinel9=0):
    """Synthetic dataset from the library's seeded generator."""
    import epigpu_amd as E
    words = E.synth_genotypes(seed, nid, nsnp, maf[0], maf[1], gmiss)
    y = E.synth_phenotype(seed, 0, nid, dist, pmiss)
    if sentinel9:
        rng = np.random.default_rng(seed)
        y[rng.choice(nid, sentinel9, replace=False)] = -9.0
    return words, y


def rel_err(x, ref):
    """|x - ref| / max(|ref|, 1): relative above 1, absolute below (tiny-F policy)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.abs(x - ref) / np.maximum(np.abs(ref), 1.0)
