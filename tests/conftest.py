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


def hits_vs_oracle(g, codes, y, starts, sizes, test, t1, t2, safe=False, tol=1e-5):
    """Checks a scan's hit list against the oracle's exhaustive scan of the same blocks: the hit sets are
    equal apart from threshold ties (pairs whose oracle F or F2 lies within tol of its threshold), and every
    reported hit's FP64 F and F2 are within tol of the oracle's. g: a dict with 'hits', 'F64', 'F2_64'.
    Returns (number of hits, listed ties)."""
    import oracle as O
    n_ref, ref = O.scan(codes, y, starts, sizes, test, t1, t2, safe=safe)
    gs = list(zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist()))
    assert len(set(gs)) == len(gs), "a pair is reported twice"
    rs = set(zip(ref["snp1"].tolist(), ref["snp2"].tolist()))
    assert n_ref == len(rs)
    if gs:
        orc = O.pairs(codes, y, np.array(gs, dtype=np.int32), safe=safe)
        ef = rel_err(g["F64"], orc["F"])
        assert ef.max() <= tol, f"reported F off the oracle's by {ef.max():.3g}"
        if test == "i":
            ef2 = rel_err(g["F2_64"], orc["F2"])
            assert ef2.max() <= tol, f"reported F2 off the oracle's by {ef2.max():.3g}"
    diff = sorted(set(gs) ^ rs)
    ties = []
    if diff:
        o = O.pairs(codes, y, np.array(diff, dtype=np.int32), safe=safe)
        for i, pr in enumerate(diff):
            near = abs(o["F"][i] - t1) <= tol * max(t1, 1.0)
            if test == "i":
                near = near or abs(o["F2"][i] - t2) <= tol * max(t2, 1.0)
            assert near, f"pair {pr} differs from the oracle (F {o['F'][i]:.9g}, F2 {o['F2'][i]:.9g})"
            ties.append(pr)
    return len(gs), ties
