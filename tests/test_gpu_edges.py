"""GPU edge cases of the scan against the FP64 oracle: phenotype scales far from 1, constant and almost
all-missing phenotypes, the smallest inputs, one-SNP chromosomes, and the 64-row MFMA tiles (512-thread
workgroups) on the per-pair dump."""
import numpy as np
import pytest

import oracle as O
from conftest import make_dataset, rel_err

pytestmark = pytest.mark.gpu

TOL_F = 1e-5
FORMS = ["plain", "mfma"]


def dump_vs_oracle(scanner, words, y, nid, form, rect=None):
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    nsnp = words.shape[0]
    a0, a1, b0, b1 = rect or (0, nsnp, 0, nsnp)
    counts, sums, F, F2 = scanner.pair_dump(a0, a1, b0, b1, test="i", form=form)
    aa, bb = np.meshgrid(np.arange(a0, a1), np.arange(b0, b1), indexing="ij")
    sel = aa > bb
    ref = O.pairs(O.unpack(words, nid), y, np.stack([aa[sel], bb[sel]], axis=1))
    assert (counts[sel] == ref["counts"]).all()
    return F[sel], F2[sel], sums[sel], ref


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("scale", [1e-30, 1e-6, 1e6, 1e30])
def test_phenotype_scale(scanner, form, scale):
    """F is scale-free: phenotypes near the ends of the FP32 range give the oracle's F (and the unit-scale F)."""
    nid, nsnp = 400, 96
    words, y = make_dataset(41, nid, nsnp, gmiss=0.01, pmiss=0.02)
    ys = (y.astype(np.float64) * scale).astype(np.float32)
    print(f"scale {scale:g} form={form}", flush=True)
    F, F2, sums, ref = dump_vs_oracle(scanner, words, ys, nid, form)
    ok = ref["df1"] >= 1
    assert rel_err(F[ok], ref["F"][ok]).max() <= TOL_F
    assert rel_err(F2[ok], ref["F2"][ok]).max() <= TOL_F
    assert np.abs(sums - ref["sums"]).max() <= 1e-5 * scale * nid


@pytest.mark.parametrize("form", FORMS)
def test_constant_phenotype_gives_no_hits(scanner, form):
    """All present phenotypes equal: SSW = 0, F is 0/0 in FP64 and no pair is a hit."""
    nid, nsnp = 300, 80
    words, _ = make_dataset(42, nid, nsnp)
    y = np.full(nid, 2.5, dtype=np.float32)
    y[::7] = np.nan
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    g = scanner.scan(test="i", form=form, t1=0.5, t2=0.5)
    assert g["nhits"] == 0
    F, F2, sums, ref = dump_vs_oracle(scanner, words, y, nid, form)
    assert np.isnan(ref["F"]).all() and np.isnan(F).all()


@pytest.mark.parametrize("form", FORMS)
def test_two_present_phenotypes(scanner, form):
    """Only two individuals with a phenotype: d2 < 1 everywhere, no hits, counts exact."""
    nid, nsnp = 200, 40
    words, _ = make_dataset(43, nid, nsnp)
    y = np.full(nid, np.nan, dtype=np.float32)
    y[3], y[150] = 1.0, -2.0
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    g = scanner.scan(test="f", form=form, t1=0.0)
    assert g["nhits"] == 0 and g["pairs"] == nsnp * (nsnp - 1) // 2
    dump_vs_oracle(scanner, words, y, nid, form)


@pytest.mark.parametrize("form", FORMS)
def test_smallest_inputs(scanner, form):
    """nid = 2 individuals and 2 SNPs: one pair, one block, each individual in its own cell (d2 = 0: no hit)."""
    import epigpu_amd as E
    codes = np.array([[0, 2], [1, 0]], dtype=np.int8)
    words = E.pack(codes)
    y = np.array([1.0, 3.0], dtype=np.float32)
    scanner.set_genotypes(words, 2)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [2])
    g = scanner.scan(test="f", form=form, t1=0.0)
    assert g["pairs"] == 1 and g["nhits"] == 0
    counts, sums, F, F2 = scanner.pair_dump(1, 2, 0, 1, test="i", form=form)
    ref = O.pairs(O.unpack(words, 2), y, np.array([[1, 0]]))
    assert counts[0, 0].tolist() == ref["counts"][0].tolist() == [0, 0, 1, 1, 0, 0, 0, 0, 0]
    assert np.allclose(sums[0, 0], ref["sums"][0])


@pytest.mark.parametrize("form", FORMS)
def test_one_snp_chromosomes(scanner, form):
    """Chromosomes of one SNP: empty diagonal blocks, block numbering as the reference."""
    nid, sizes = 350, [1, 30, 1, 25]
    nsnp = sum(sizes)
    words, y = make_dataset(44, nid, nsnp, gmiss=0.01, pmiss=0.02)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes(starts, sizes)
    g = scanner.scan(test="i", form=form, t1=1.0, t2=1.0)
    n_ref, ref = O.scan(O.unpack(words, nid), y, starts, sizes, "i", 1.0, 1.0)
    gs = set(zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist()))
    rs = set(zip(ref["snp1"].tolist(), ref["snp2"].tolist()))
    for pr in gs ^ rs:
        r = ref[(ref["snp1"] == pr[0]) & (ref["snp2"] == pr[1])]
        assert len(r) == 0 or min(abs(r["F"][0] - 1.0), abs(r["F2"][0] - 1.0)) < 1e-5
    assert g["pairs"] == nsnp * (nsnp - 1) // 2
    rb = {(int(a), int(c)): int(k) for a, c, k in zip(ref["snp1"], ref["snp2"], ref["block"])}
    for a, c, k in zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist(), g["block"].tolist()):
        if (a, c) in rb:
            assert rb[(a, c)] == k
    assert (np.diff(g["block"]) >= 0).all()


@pytest.mark.parametrize("nid", [300, 1000])
def test_mfma_64_row_tiles_dump(scanner, nid):
    """The 64-row FAST tiles and 64 x 32 GENERAL tiles (512-thread workgroups) on the per-pair dump; with
    0.4 % missing calls over 150 SNPs both tile kinds occur."""
    nsnp = 150
    words, y = make_dataset(45 + nid, nid, nsnp, gmiss=0.004, pmiss=0.02)
    default = scanner.schedule("mfma")
    try:
        scanner.set_schedule("mfma", fast_rows=64)
        print(f"mfma 64-row dump nid={nid}", flush=True)
        F, F2, sums, ref = dump_vs_oracle(scanner, words, y, nid, "mfma")
        ok = ref["df1"] >= 1
        assert rel_err(F[ok], ref["F"][ok]).max() <= TOL_F
        assert rel_err(F2[ok], ref["F2"][ok]).max() <= TOL_F
    finally:
        scanner.set_schedule("mfma", **default)


def test_single_scan_beyond_one_launch(scanner):
    """More tiles than one launch may cover (HIP counts a launch's work-items in 32 bits): 200,000 SNPs with
    missing calls give about 9.8 million 64 x 32 GENERAL tiles of 512 threads, 5e9 work-items. The single
    scan must split its launches and find exactly the hits of the persistent batch launch of the same data."""
    import epigpu_amd as E
    nid, nsnp = 100, 200000
    words, y = make_dataset(0xB16, nid, nsnp, maf=(0.1, 0.5), gmiss=0.001)
    default = scanner.schedule("mfma")
    try:
        scanner.set_genotypes(words, nid)
        scanner.set_phenotypes(y)
        scanner.set_chromosomes([0], [nsnp])
        scanner.set_schedule("mfma", fast_rows=64, super_a=16, super_b=2)
        print("single scan beyond one launch", flush=True)
        one = scanner.scan(test="f", form="mfma", t1=7.5)
        assert one["launches"] >= 2 and not one["overflow"]
        assert one["pairs"] == nsnp * (nsnp - 1) // 2
        b = E.Batch(scanner)
        b.add_phenotypes(b.add_dataset(words, nid, [0], [nsnp]), y)
        b.set_schedule("mfma", fast_rows=64, super_a=16, super_b=2)
        st = b.scan(test="f", form="mfma", t1=7.5)
        g = b.hits(0)
        b.close()
        assert not st["overflow"] and g["nhits"] == one["nhits"] > 100
        assert (g["hits"]["snp1"] == one["hits"]["snp1"]).all() and (g["hits"]["snp2"] == one["hits"]["snp2"]).all()
    finally:
        scanner.set_schedule("mfma", **default)


@pytest.mark.parametrize("form,dense", [("plain", 0), ("mfma", 0), ("mfma", 1)])
def test_scan_in_pieces_matches_one_list(scanner, form, dense):
    """A tile list above the context's bound is built, uploaded and launched in pieces (whole rows of
    super-tiles, across chromosome blocks). With the bound cut to 7, 60 and 500 tiles the scan must find exactly
    the hits and pair count of the one-list scan, for the whole range and for a block sub-range; afterwards the
    default bound gives the one-list scan again. Some SNPs have missing calls, so FAST tiles mix with SPARSE
    tiles (MFMA form, dense_missing 0) or GENERAL tiles (dense_missing 1, and the plain form)."""
    import epigpu_amd as E
    nid, nsnp = 300, 1500
    words, y = make_dataset(0x91E, nid, nsnp, maf=(0.05, 0.5), gmiss=0.0)
    codes = E.unpack(words, nid)
    codes[np.random.default_rng(3).random(nsnp) < 0.1, 7] = 3  # individual 7 misses a call on a tenth of the SNPs
    words = E.pack(codes)
    starts, sizes = [0, 700, 700, 1200], [700, 0, 500, 300]
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes(starts, sizes)
    scanner.set_hit_capacity(1 << 22)
    default = scanner.schedule("mfma")
    scanner.set_schedule("mfma", dense_missing=dense)
    key = lambda r: (r["hits"]["snp1"], r["hits"]["snp2"], r["block"])
    try:
        for bb, be in [(0, 0), (2, 8)]:
            one = scanner.scan(test="i", form=form, t1=2.0, t2=2.5, blk_begin=bb, blk_end=be)
            assert one["nhits"] > 1000 and not one["overflow"]
            tiles = scanner.last_tiles()
            if form == "mfma":
                assert tiles["sparse" if dense == 0 else "general"] > 0, tiles
            ntiles = sum(tiles.values())
            for cap in (7, 60, 500):
                print(f"pieces form={form} dense={dense} blocks=[{bb},{be}) max_tiles={cap} of {ntiles}", flush=True)
                scanner.set_max_tiles(cap)
                r = scanner.scan(test="i", form=form, t1=2.0, t2=2.5, blk_begin=bb, blk_end=be)
                assert r["pairs"] == one["pairs"] and r["nhits"] == one["nhits"] and not r["overflow"]
                if ntiles > cap:
                    assert r["launches"] > one["launches"]
                for x, z in zip(key(r), key(one)):
                    assert (x == z).all()
                assert np.array_equal(r["F64"], one["F64"]) and np.array_equal(r["F2_64"], one["F2_64"])
            scanner.set_max_tiles(1 << 25)
            again = scanner.scan(test="i", form=form, t1=2.0, t2=2.5, blk_begin=bb, blk_end=be)
            assert again["launches"] == one["launches"] and again["nhits"] == one["nhits"]
    finally:
        scanner.set_max_tiles(1 << 25)
        scanner.set_schedule("mfma", **default)


@pytest.mark.parametrize("gmiss", [0.0, 0.001])
def test_largest_cohort(scanner, gmiss):
    """nid = 2,096,896, the largest the transcode grid takes (65,528 rows of 32 individuals): 16,382 K-loop
    stages per tile, limb sums up to 2.7e8 in their 32-bit accumulators. MFMA tables exact and F / F2 within
    1e-5 of the oracle; the plain form's counts exact (its FP32 sums over 2 million values are not held to
    1e-5). One K-pad more is refused with EPI_ERR_RANGE before anything changes."""
    import epigpu_amd as E
    nid, nsnp = 2096896, 40
    words, y = make_dataset(77, nid, nsnp, gmiss=gmiss, pmiss=0.01)
    print(f"largest cohort nid={nid} gmiss={gmiss}", flush=True)
    F, F2, sums, ref = dump_vs_oracle(scanner, words, y, nid, "mfma")
    ok = ref["df1"] >= 1
    assert rel_err(F[ok], ref["F"][ok]).max() <= TOL_F
    assert rel_err(F2[ok], ref["F2"][ok]).max() <= TOL_F
    dump_vs_oracle(scanner, words, y, nid, "plain")
    big = nid + 256
    with pytest.raises(E.EpiError) as e:
        scanner.set_genotypes(np.zeros((2, E.layout(big)[2]), np.int32), big)
    assert e.value.code == -6
    with pytest.raises(E.EpiError) as e:
        scanner.set_phenotypes(np.zeros(big, np.float32))
    assert e.value.code == -6
    counts, _, _, _ = scanner.pair_dump(1, 2, 0, 1, form="mfma")
    assert counts[0, 0].sum() == ref["counts"][0].sum()


def test_pieces_at_default_bound(scanner):
    """400,000 SNPs with missing calls under the default schedule: 78 million 32 x 32 GENERAL tiles, which the
    default bound of 2^25 tiles splits into three pieces. Same hits, F64 values and pair count as one 78-million-
    tile list."""
    nid, nsnp = 100, 400000
    words, y = make_dataset(0x400, nid, nsnp, maf=(0.1, 0.5), gmiss=0.001)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    try:
        print("pieces at the default bound", flush=True)
        r = scanner.scan(test="i", form="mfma", t1=9.0, t2=14.0)
        scanner.set_max_tiles((1 << 30) - 1)
        print("one list", flush=True)
        one = scanner.scan(test="i", form="mfma", t1=9.0, t2=14.0)
    finally:
        scanner.set_max_tiles(1 << 25)
    assert r["pairs"] == one["pairs"] == nsnp * (nsnp - 1) // 2
    assert r["nhits"] == one["nhits"] > 100 and not r["overflow"] and not one["overflow"]
    assert r["launches"] > one["launches"]
    assert (r["hits"]["snp1"] == one["hits"]["snp1"]).all() and (r["hits"]["snp2"] == one["hits"]["snp2"]).all()
    assert np.array_equal(r["F64"], one["F64"]) and np.array_equal(r["F2_64"], one["F2_64"])
