"""SPARSE tiles of the MFMA form (tiles with missing calls whose SNPs each miss at most max(8, nid / 16) calls):
the FAST contraction plus per-pair sums over the individuals that miss a call.

Most tests force SPARSE tiles wherever they are allowed (dense_missing = 2), so that missing rates above the
default's 0.67 % run through them too. Checks: per-pair tables bit-exact and F / F2 within 1e-5 of the oracle, and
cell sums and F identical to the GENERAL tiles' (dense_missing = 1) on the same data, since both end in the same
exact integer sums; one dataset whose scan has FAST, SPARSE and GENERAL tiles; the default's choice by missing
rate; outliers and safe mode; whole scans' hit lists against the oracle on chromosomes whose sizes are not
multiples of 64; the batch against single scans.
"""
import numpy as np
import pytest

import oracle as O
from conftest import hits_vs_oracle, make_dataset, rel_err

pytestmark = pytest.mark.gpu

TOL_F = 1e-5


@pytest.fixture
def sch(scanner):
    """The session scanner; its MFMA schedule is restored afterwards."""
    saved = scanner.schedule("mfma")
    yield scanner
    scanner.set_schedule("mfma", **saved)


def oracle_rect(codes, y, n, safe=False):
    aa, bb = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    sel = aa > bb
    pr = np.stack([aa[sel], bb[sel]], axis=1)
    return sel, O.pairs(codes, y, pr, safe=safe)


def dump_kinds(s, words, y, nid, nsnp, rows, nine=False, safe=False):
    """Dumps of every pair with SPARSE tiles wherever allowed (dense_missing 2) and without (1), each with its
    tile counts, checked against the oracle; returns both."""
    s.set_genotypes(words, nid)
    s.set_phenotypes(y, nine_is_value=nine)
    codes = O.unpack(words, nid)
    sel, ref = oracle_rect(codes, y, nsnp, safe=safe)
    ok = ref["df1"] >= 1
    out = {}
    for dense in (2, 1):
        s.set_schedule("mfma", fast_rows=rows, dense_missing=dense)
        print(f"sparse dump rows={rows} dense={dense} nid={nid} nsnp={nsnp}", flush=True)
        counts, sums, F, F2 = s.pair_dump(0, nsnp, 0, nsnp, test="i", form="mfma")
        tiles = s.last_tiles()
        assert (counts[sel] == ref["counts"]).all(), f"dense={dense}: counts differ"
        assert rel_err(F[sel][ok], ref["F"][ok]).max() <= TOL_F, f"dense={dense}: F off"
        assert rel_err(F2[sel][ok], ref["F2"][ok]).max() <= TOL_F, f"dense={dense}: F2 off"
        out[dense] = (counts[sel], sums[sel], F[sel], F2[sel], tiles)
    # the same exact integer sums, whichever tile kind produced them
    assert np.array_equal(out[2][1], out[1][1]), "cell sums differ between SPARSE and GENERAL tiles"
    assert np.array_equal(out[2][2], out[1][2]) and np.array_equal(out[2][3], out[1][3])
    return out


@pytest.mark.parametrize("rows", [32, 64])
@pytest.mark.parametrize("gmiss", [0.01, 0.04])
def test_sparse_dump_equals_general_and_oracle(sch, rows, gmiss):
    nid, nsnp = 1000, 200
    words, y = make_dataset(0x5A1 + int(gmiss * 100), nid, nsnp, maf=(0.01, 0.5), gmiss=gmiss, pmiss=0.03,
                            sentinel9=5)
    out = dump_kinds(sch, words, y, nid, nsnp, rows)
    t0, t1 = out[2][4], out[1][4]
    assert t0["sparse"] > 0 and t0["general"] == 0, t0
    assert t1["sparse"] == 0 and t1["general"] > 0, t1


@pytest.mark.parametrize("gmiss,kind", [(0.003, "sparse"), (0.02, "general")])
def test_default_picks_tile_kind_by_missing_rate(sch, gmiss, kind):
    """The default (dense_missing 0) takes SPARSE tiles at 0.3 % missing calls and GENERAL tiles at 2 %; the
    pairs match the oracle either way."""
    nid, nsnp = 2000, 192
    words, y = make_dataset(0xDEF + int(gmiss * 1000), nid, nsnp, maf=(0.02, 0.5), gmiss=gmiss, pmiss=0.02)
    sch.set_genotypes(words, nid)
    sch.set_phenotypes(y)
    sch.set_schedule("mfma", fast_rows=64, dense_missing=0)
    print(f"default tile kind gmiss={gmiss}", flush=True)
    counts, sums, F, F2 = sch.pair_dump(0, nsnp, 0, nsnp, test="i", form="mfma")
    t = sch.last_tiles()
    other = "general" if kind == "sparse" else "sparse"
    assert t[kind] > 0 and t[other] == 0 and t["fast"] == 0, t
    sel, ref = oracle_rect(O.unpack(words, nid), y, nsnp)
    assert (counts[sel] == ref["counts"]).all()
    ok = ref["df1"] >= 1
    assert rel_err(F[sel][ok], ref["F"][ok]).max() <= TOL_F


def test_sparse_fast_and_general_in_one_scan(sch):
    """SNPs 0-63 without missing calls, 64-191 with about 1 %, 192-255 with some SNPs missing 20 % of their
    calls (above nid / 16): 64-row tiles of all three kinds."""
    nid, nsnp = 1000, 256
    words, y = make_dataset(0x3C1D, nid, nsnp, maf=(0.02, 0.5), gmiss=0.01, pmiss=0.02)
    import epigpu_amd as E
    codes = O.unpack(words, nid)
    rng = np.random.default_rng(5)
    c0 = codes[:64]
    c0[c0 == 3] = 0
    for snp in (200, 231, 250):
        codes[snp, rng.random(nid) < 0.2] = 3
    words = E.pack(codes)
    out = dump_kinds(sch, words, y, nid, nsnp, 64)
    t0 = out[2][4]
    assert t0["fast"] > 0 and t0["sparse"] > 0 and t0["general"] > 0, t0


@pytest.mark.parametrize("rows", [32, 64])
def test_sparse_outliers(sch, rows):
    """Far values carried exactly (outliers) on individuals that miss calls in some SNPs."""
    nid, nsnp = 1200, 140
    words, y = make_dataset(0x0F1, nid, nsnp, maf=(0.02, 0.5), gmiss=0.02, pmiss=0.02)
    y = y.copy()
    y[[3, 500, 901]] = [400.0, -2500.0, 1e4]
    out = dump_kinds(sch, words, y, nid, nsnp, rows)
    assert out[2][4]["sparse"] > 0
    from epigpu_amd import scan as S
    assert S.phen_encoding(y)["outliers"] == 3


def test_sparse_safe_mode(sch):
    """-9 as a phenotype value (safe mode) with missing calls."""
    nid, nsnp = 900, 130
    words, y = make_dataset(0x5AF, nid, nsnp, maf=(0.02, 0.5), gmiss=0.015, pmiss=0.02, sentinel9=30)
    out = dump_kinds(sch, words, y, nid, nsnp, 64, nine=True, safe=True)
    assert out[2][4]["sparse"] > 0


@pytest.mark.parametrize("rows", [32, 64])
def test_sparse_scan_hits(sch, rows):
    """Whole scans over two chromosomes of 333 and 290 SNPs (the virtual SNP order pads each to 64) against the
    oracle's hit list, and the same hits as GENERAL tiles."""
    nid, sizes = 1200, [333, 290]
    nsnp = sum(sizes)
    starts = [0, sizes[0]]
    words, y = make_dataset(0xC2C2, nid, nsnp, maf=(0.01, 0.5), gmiss=0.01, pmiss=0.02)
    sch.set_genotypes(words, nid)
    sch.set_phenotypes(y)
    sch.set_chromosomes(starts, sizes)
    codes = O.unpack(words, nid)
    res = {}
    for dense in (2, 1):
        sch.set_schedule("mfma", fast_rows=rows, dense_missing=dense)
        print(f"sparse scan rows={rows} dense={dense}", flush=True)
        g = sch.scan(test="i", form="mfma", t1=2.2, t2=2.6)
        tiles = sch.last_tiles()
        assert (tiles["sparse"] > 0) == (dense == 2) and tiles["fast"] == 0, tiles
        n, ties = hits_vs_oracle(g, codes, y, starts, sizes, "i", 2.2, 2.6)
        assert n > 200
        res[dense] = set(zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist()))
    assert res[2] == res[1]


@pytest.mark.parametrize("rows", [32, 64])
def test_sparse_batch_equals_single_scans(sch, rows):
    """Two datasets with missing calls and two phenotypes each in one batch, SPARSE tiles in the persistent
    kernel, against single scans of the same data."""
    import epigpu_amd as E
    data = []
    for k, (nid, nsnp) in enumerate([(700, 300), (1500, 180)]):
        words, y = make_dataset(0xBA7 + k, nid, nsnp, maf=(0.01, 0.5), gmiss=0.012, pmiss=0.02)
        y2 = E.synth_phenotype(0xBA7 + k, 1, nid, 1, 0.02)
        data.append((words, nid, nsnp, np.stack([y, y2])))
    with E.Batch(sch) as b:
        b.set_schedule("mfma", fast_rows=rows, dense_missing=2)
        first = []
        for words, nid, nsnp, ys in data:
            d = b.add_dataset(words, nid, [0], [nsnp])
            first.append(b.add_phenotypes(d, ys))
        print(f"sparse batch rows={rows}", flush=True)
        b.scan(test="i", form="mfma", t1=2.4, t2=2.8)
        got = {}
        for ids in first:
            for sid in ids:
                h = b.hits(sid)
                got[sid] = set(zip(h["hits"]["snp1"].tolist(), h["hits"]["snp2"].tolist()))
    sch.set_schedule("mfma", fast_rows=rows, dense_missing=2)
    for (words, nid, nsnp, ys), ids in zip(data, first):
        sch.set_genotypes(words, nid)
        sch.set_phenotypes(ys)
        sch.set_chromosomes([0], [nsnp])
        for r, sid in enumerate(ids):
            g = sch.scan(test="i", form="mfma", t1=2.4, t2=2.8, phen=r)
            assert sch.last_tiles()["sparse"] > 0
            assert set(zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist())) == got[sid]
            assert len(got[sid]) > 50


@pytest.mark.parametrize("rows", [32, 64])
def test_sparse_long_miss_lists(sch, rows):
    """N = 20,001 at 0.5 % missing calls: about 100 individuals per SNP's miss list, so the passes load each
    list in several chunks of 64 (and lists of every length around the chunk boundary occur)."""
    nid, nsnp = 20001, 130
    words, y = make_dataset(0x10C5 + rows, nid, nsnp, maf=(0.02, 0.5), gmiss=0.005, pmiss=0.02)
    import epigpu_amd as E
    m = (E.unpack(words, nid) == 3).sum(1)
    assert m.max() > 64 and m.min() < 128
    out = dump_kinds(sch, words, y, nid, nsnp, rows)
    assert out[2][4]["sparse"] > 0
