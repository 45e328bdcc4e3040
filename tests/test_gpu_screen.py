"""The FP32 screen of FAST tiles against the oracle's exhaustive scan.

A FAST tile screens each pair before the FP64 decision: 64-row tiles with the all-occupied form when every pair
of a wave has nine occupied cells, else the general form; 32-row tiles with the per-lane form. Each must keep
every pair that the FP64 decision makes a hit. The datasets have no missing call (all tiles FAST) and thresholds
low enough for hundreds of hits: a small cohort with rare SNPs (empty cells common, so the nfac rules and the
empty-cell terms matter) and a larger one with common SNPs (nine occupied cells nearly everywhere). Checks:
hit lists equal to the oracle's apart from threshold ties, with F / F2 within 1e-5.
"""
import numpy as np
import pytest

import oracle as O
from conftest import hits_vs_oracle, make_dataset

pytestmark = pytest.mark.gpu

COHORTS = {
    "small_rare": dict(nid=400, nsnp=300, maf=(0.01, 0.5)),
    "large_common": dict(nid=3000, nsnp=200, maf=(0.25, 0.5)),
}


@pytest.fixture
def sch(scanner):
    """The session scanner; its MFMA schedule is restored afterwards."""
    saved = scanner.schedule("mfma")
    yield scanner
    scanner.set_schedule("mfma", **saved)


@pytest.mark.parametrize("rows", [32, 64])
@pytest.mark.parametrize("cohort", sorted(COHORTS))
@pytest.mark.parametrize("test", ["i", "f"])
def test_fast_screen_hits_equal_oracle(sch, rows, cohort, test):
    c = COHORTS[cohort]
    words, y = make_dataset(0x5C4EE + c["nid"] + rows, c["nid"], c["nsnp"], maf=c["maf"], pmiss=0.02)
    sch.set_genotypes(words, c["nid"])
    sch.set_phenotypes(y)
    sch.set_chromosomes([0], [c["nsnp"]])
    sch.set_schedule("mfma", fast_rows=rows)
    t1, t2 = (2.2, 2.6) if test == "i" else (2.0, 10.5)
    print(f"screen rows={rows} cohort={cohort} test={test}", flush=True)
    g = sch.scan(test=test, form="mfma", t1=t1, t2=t2)
    t = sch.last_tiles()
    assert t["fast"] > 0 and t["sparse"] == 0 and t["general"] == 0, t
    n, ties = hits_vs_oracle(g, O.unpack(words, c["nid"]), y, [0], [c["nsnp"]], test, t1, t2)
    assert n > 100


@pytest.mark.parametrize("rows", [32, 64])
def test_fast_screen_safe_mode(sch, rows):
    """Safe mode (no nfac > 6 rule for test i) on the small cohort, where pairs with empty cells abound."""
    c = COHORTS["small_rare"]
    words, y = make_dataset(0x5AFE + rows, c["nid"], c["nsnp"], maf=c["maf"], pmiss=0.02)
    sch.set_genotypes(words, c["nid"])
    sch.set_phenotypes(y)
    sch.set_chromosomes([0], [c["nsnp"]])
    sch.set_schedule("mfma", fast_rows=rows)
    print(f"screen safe rows={rows}", flush=True)
    g = sch.scan(test="i", form="mfma", t1=2.2, t2=2.6, safe=True)
    n, ties = hits_vs_oracle(g, O.unpack(words, c["nid"]), y, [0], [c["nsnp"]], "i", 2.2, 2.6, safe=True)
    assert n > 100
    # the safe rule admits pairs with at most six occupied cells, which the default rule drops
    g0 = sch.scan(test="i", form="mfma", t1=2.2, t2=2.6)
    assert len(g["hits"]) > len(g0["hits"])
