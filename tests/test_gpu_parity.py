"""GPU parity: pair tables and F statistics of both kernel forms against the FP64 oracle.

Bar: cell counts bit-exact; F and F2 within 1e-5 of the oracle, relative for
values >= 1 and absolute below 1 (rel_err); MFMA cell sums within the fixed-
point quantisation of the phenotype, plain cell sums within FP32 accumulation.
"""
import numpy as np
import pytest

import oracle as O
from conftest import make_dataset, rel_err

pytestmark = pytest.mark.gpu

TOL_F = 1e-5


def oracle_rect(codes, y, a0, a1, b0, b1):
    aa, bb = np.meshgrid(np.arange(a0, a1), np.arange(b0, b1), indexing="ij")
    sel = aa > bb
    pr = np.stack([aa[sel], bb[sel]], axis=1)
    return sel, O.pairs(codes, y, pr)


def check_rect(scanner, words, y, nid, a0, a1, b0, b1, form, test="i"):
    import epigpu_amd as E
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    counts, sums, F, F2 = scanner.pair_dump(a0, a1, b0, b1, test=test, form=form)
    codes = O.unpack(words, nid)
    sel, ref = oracle_rect(codes, y, a0, a1, b0, b1)
    c = counts[sel]
    assert (c == ref["counts"]).all(), f"{form}: counts differ in {(c != ref['counts']).any(1).sum()} pairs"
    ok = ref["df1"] >= 1
    ef = rel_err(F[sel][ok], ref["F"][ok])
    ef2 = rel_err(F2[sel][ok], ref["F2"][ok])
    assert ef.max() <= TOL_F, f"{form}: max F error {ef.max():.3g}"
    assert ef2.max() <= TOL_F, f"{form}: max F2 error {ef2.max():.3g}"
    s = sums[sel]
    scale = np.abs(y[np.isfinite(y) & (y != -9)]).max()
    if form == "mfma":
        tol = (ref["counts"] + 1) * scale * 2.0 ** -22
    else:
        # derived cells inherit the FP32 accumulation error of the class totals
        tol = 1e-7 * nid * scale + 1e-9
    assert (np.abs(s - ref["sums"]) <= tol).all(), f"{form}: cell sums off by {np.abs(s - ref['sums']).max():.3g}"
    return ef.max(), ef2.max()


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_rect_no_missing(scanner, form):
    nid, nsnp = 500, 260
    words, y = make_dataset(11, nid, nsnp)
    print(f"rect no-missing form={form} nid={nid} nsnp={nsnp}", flush=True)
    check_rect(scanner, words, y, nid, 0, nsnp, 0, nsnp, form)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_rect_missing(scanner, form):
    nid, nsnp = 777, 200
    words, y = make_dataset(12, nid, nsnp, maf=(0.01, 0.5), gmiss=0.03, pmiss=0.05, sentinel9=7)
    print(f"rect missing form={form} nid={nid} nsnp={nsnp}", flush=True)
    check_rect(scanner, words, y, nid, 0, nsnp, 0, nsnp, form)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("nid", [16, 33, 257, 1000])
def test_rect_ragged_nid(scanner, form, nid):
    nsnp = 70
    words, y = make_dataset(13 + nid, nid, nsnp, maf=(0.05, 0.5), gmiss=0.02 if nid > 100 else 0.0, pmiss=0.02)
    print(f"rect ragged form={form} nid={nid}", flush=True)
    check_rect(scanner, words, y, nid, 0, nsnp, 0, nsnp, form)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("gmiss", [0.0, 0.01])
@pytest.mark.parametrize("nid", [20001, 100003])
def test_rect_large_nid(scanner, form, gmiss, nid):
    """N = 20,001 (c5's scale) and 100,003 (odd numbers of 128-individual stages): FAST tiles without missing
    calls, GENERAL tiles with them; counts exact and F / F2 within tolerance at long K loops."""
    nsnp = 72
    words, y = make_dataset(0x5EED + int(gmiss * 100), nid, nsnp, maf=(0.01, 0.5), gmiss=gmiss, pmiss=0.01)
    print(f"rect large form={form} nid={nid} gmiss={gmiss}", flush=True)
    check_rect(scanner, words, y, nid, 0, nsnp, 0, nsnp, form)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_rect_offset_window(scanner, form):
    """A rectangle away from the origin and not tile aligned."""
    nid, nsnp = 640, 400
    words, y = make_dataset(14, nid, nsnp, dist=2)
    print(f"rect offset form={form}", flush=True)
    check_rect(scanner, words, y, nid, 137, 331, 29, 250, form)


def hit_set_check(g, n_ref, ref, test, t1, t2, tol=1e-5):
    """Hit sets equal apart from pairs whose FP64 F or F2 lies within tol of a threshold."""
    gs = {(int(a), int(b)) for a, b in zip(g["hits"]["snp1"], g["hits"]["snp2"])}
    rs = {(int(a), int(b)) for a, b in zip(ref["snp1"], ref["snp2"])}
    assert g["nhits"] == len(gs) and n_ref == len(rs)
    diff = gs ^ rs
    return diff


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_scan_c1_config(scanner, form):
    """Config c1: N=500, M=2000, 8df test at p < 1e-4 (t1 = 4.084638)."""
    import epigpu_amd as E
    nid, nsnp = 500, 2000
    words, y = make_dataset(0xE9150001, nid, nsnp)
    t1 = float(np.float32(E.f_isf(1e-4, 8, nid - 9)))
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    print(f"scan c1 form={form}", flush=True)
    g = scanner.scan(test="f", form=form, t1=t1)
    codes = O.unpack(words, nid)
    n_ref, ref = O.scan(codes, y, [0], [nsnp], "f", t1)
    diff = hit_set_check(g, n_ref, ref, "f", t1, 10.5)
    assert not diff, f"hit sets differ: {sorted(diff)[:10]}"
    assert (g["hits"]["snp1"] == ref["snp1"]).all() and (g["hits"]["snp2"] == ref["snp2"]).all()
    assert (g["hits"]["df1"] == ref["df1"]).all() and (g["hits"]["df2"] == ref["df2"]).all()
    assert rel_err(g["F64"], ref["F"]).max() <= TOL_F
    assert g["pairs"] == nsnp * (nsnp - 1) // 2


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_scan_two_chromosomes_missing(scanner, form):
    """c2-shaped (scaled down): 2 chromosomes, 1% missing calls, 2% NaN phenotypes, 8df+4df."""
    nid, sizes = 1000, [700, 500]
    nsnp = sum(sizes)
    words, y = make_dataset(0xE9150002, nid, nsnp, maf=(0.01, 0.5), gmiss=0.01, pmiss=0.02)
    t1, t2 = 3.0, 4.0  # low thresholds: many hits exercise the compaction
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0, 700], sizes)
    print(f"scan 2chr form={form}", flush=True)
    g = scanner.scan(test="i", form=form, t1=t1, t2=t2)
    codes = O.unpack(words, nid)
    n_ref, ref = O.scan(codes, y, [0, 700], sizes, "i", t1, t2)
    assert n_ref > 1000
    gs = {(int(a), int(b)): (f, f2) for a, b, f, f2 in zip(g["hits"]["snp1"], g["hits"]["snp2"], g["F64"], g["F2_64"])}
    rs = {(int(a), int(b)): (f, f2) for a, b, f, f2 in zip(ref["snp1"], ref["snp2"], ref["F"], ref["F2"])}
    diff = set(gs) ^ set(rs)
    for pr in diff:  # only near-threshold ties may differ
        f, f2 = rs.get(pr, gs.get(pr))
        assert abs(f - t1) <= 1e-5 * t1 or abs(f2 - t2) <= 1e-5 * t2, f"non-tie difference {pr}"
    common = sorted(set(gs) & set(rs))
    ef = max(rel_err(gs[p][0], rs[p][0]) for p in common)
    ef2 = max(rel_err(gs[p][1], rs[p][1]) for p in common)
    assert ef <= TOL_F and ef2 <= TOL_F
    # block labels follow the reference order: (0,0)=0, (1,0)=1, (1,1)=2
    blk = g["block"]
    a, b = g["hits"]["snp1"], g["hits"]["snp2"]
    exp = np.where(b >= 700, 2, np.where(a >= 700, 1, 0))
    assert (blk == exp).all()
    assert (np.diff(blk) >= 0).all()


def test_scan_block_range(scanner):
    """A scan restricted to one block equals the full scan's hits of that block."""
    nid, sizes = 600, [300, 250, 200]
    nsnp = sum(sizes)
    words, y = make_dataset(77, nid, nsnp, gmiss=0.005)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0, 300, 550], sizes)
    full = scanner.scan(test="i", form="mfma", t1=3.5, t2=4.5)
    for blk in range(6):
        part = scanner.scan(test="i", form="mfma", t1=3.5, t2=4.5, blk_begin=blk, blk_end=blk + 1)
        sel = full["block"] == blk
        assert (part["hits"] == full["hits"][sel]).all()
        assert (part["block"] == blk).all()


def test_forms_agree_on_hits(scanner):
    """Plain and MFMA forms give the same hits on the same data."""
    nid, nsnp = 2000, 1500
    words, y = make_dataset(99, nid, nsnp, maf=(0.01, 0.5), gmiss=0.01, dist=2, pmiss=0.02)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    a = scanner.scan(test="i", form="plain", t1=4.0, t2=5.0)
    b = scanner.scan(test="i", form="mfma", t1=4.0, t2=5.0)
    sa = set(zip(a["hits"]["snp1"].tolist(), a["hits"]["snp2"].tolist()))
    sb = set(zip(b["hits"]["snp1"].tolist(), b["hits"]["snp2"].tolist()))
    assert len(sa ^ sb) <= 2
    assert len(sa) > 100


def test_hit_overflow_flag(scanner):
    """A tiny capacity reports the true count and the overflow flag."""
    nid, nsnp = 300, 400
    words, y = make_dataset(5, nid, nsnp)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    full = scanner.scan(test="f", form="mfma", t1=1.5)
    scanner.set_hit_capacity(100)
    small = scanner.scan(test="f", form="mfma", t1=1.5)
    scanner.set_hit_capacity(1 << 22)
    assert full["nhits"] > 1000
    assert small["nhits"] == full["nhits"] and small["overflow"]
    assert len(small["hits"]) == 100


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_safe_mode_scan(scanner, form):
    """The reference's safe mode (-s 1): -9 is a phenotype value and test 'i' has no nfac > 6 condition."""
    nid, sizes = 700, [150, 110]
    nsnp = sum(sizes)
    words, y = make_dataset(0x5AFE, nid, nsnp, maf=(0.01, 0.3), gmiss=0.01, pmiss=0.01, sentinel9=40)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y, nine_is_value=True)
    scanner.set_chromosomes([0, 150], sizes)
    t1, t2 = 3.0, 4.0
    print(f"safe scan form={form}", flush=True)
    g = scanner.scan(test="i", form=form, t1=t1, t2=t2, safe=True)
    codes = O.unpack(words, nid)
    n_ref, ref = O.scan(codes, y, [0, 150], sizes, "i", t1, t2, safe=True)
    n_norm, _ = O.scan(codes, y, [0, 150], sizes, "i", t1, t2)
    assert n_ref != n_norm  # the two modes differ on this data
    diff = hit_set_check(g, n_ref, ref, "i", t1, t2)
    for pr in diff:
        r = ref[(ref["snp1"] == pr[0]) & (ref["snp2"] == pr[1])]
        assert len(r) == 0 or min(abs(r["F"][0] - t1) / t1, abs(r["F2"][0] - t2) / t2) < 1e-5, pr
    rs = {(int(a), int(b)): (f, f2) for a, b, f, f2 in zip(ref["snp1"], ref["snp2"], ref["F"], ref["F2"])}
    for a, b, f, f2 in zip(g["hits"]["snp1"], g["hits"]["snp2"], g["F64"], g["F2_64"]):
        if (int(a), int(b)) in rs:
            assert rel_err(f, rs[(int(a), int(b))][0]) <= TOL_F and rel_err(f2, rs[(int(a), int(b))][1]) <= TOL_F
    # the per-pair tables with -9 as a value
    counts, sums, F, F2 = scanner.pair_dump(0, 60, 0, 60, test="i", form=form)
    sel, orc = oracle_rect(codes, y, 0, 60, 0, 60)
    orc = O.pairs(codes, y, np.stack(np.nonzero(sel), 1), safe=True)
    assert (counts[sel] == orc["counts"]).all()
    scanner.set_phenotypes(y)  # back to normal -9 handling for later tests


@pytest.mark.parametrize("form,gmiss", [("plain", 0.01), ("mfma", 0.0), ("mfma", 0.01)])
def test_class_relabel_invariance_full_n(scanner, form, gmiss):
    """At N = 8,000 (the 50,000-SNP workload's cohort), swapping classes 0 and 2 on a third of the SNPs (what a
    different first-seen homozygote does) moves each pair's cells to (s_a(g1), s_b(g2)) bit-exactly and leaves F
    and F2 unchanged: the 8df F is a one-way ANOVA over cells and the interaction SS sums over rows and columns.
    A kernel that mixes up class planes, weight rows or the derived class-2 cell breaks the cell map. MFMA runs
    both tile kinds (no missing calls: 64-row FAST tiles; 1 % missing: GENERAL tiles)."""
    import epigpu_amd as E
    nid, nsnp = 8000, 640
    words, y = make_dataset(41, nid, nsnp, maf=(0.02, 0.5), gmiss=gmiss, pmiss=0.01)
    codes = E.unpack(words, nid)
    flip = np.random.default_rng(5).random(nsnp) < 1 / 3
    hom = (codes == 0) | (codes == 2)
    sw = np.where(flip[:, None] & hom, 2 - codes, codes).astype(np.int8)
    print(f"relabel form={form} gmiss={gmiss} nid={nid} nsnp={nsnp} flipped={int(flip.sum())}", flush=True)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    c0, s0, F0, G0 = scanner.pair_dump(0, nsnp, 0, nsnp, test="i", form=form)
    scanner.set_genotypes(E.pack(sw), nid)
    c1, s1, F1, G1 = scanner.pair_dump(0, nsnp, 0, nsnp, test="i", form=form)
    aa, bb = np.meshgrid(np.arange(nsnp), np.arange(nsnp), indexing="ij")
    sel = aa > bb
    g1, g2 = np.arange(9) // 3, np.arange(9) % 3
    fa, fb = flip[aa[sel]][:, None], flip[bb[sel]][:, None]
    dst = 3 * np.where(fa, 2 - g1, g1) + np.where(fb, 2 - g2, g2)
    rows = np.arange(dst.shape[0])[:, None]
    mapped = np.empty_like(c0[sel])
    mapped[rows, dst] = c0[sel]
    assert (c1[sel] == mapped).all(), f"{form}: cell map broken in {(c1[sel] != mapped).any(1).sum()} pairs"
    msum = np.empty_like(s0[sel])
    msum[rows, dst] = s0[sel]
    scale = np.abs(y[np.isfinite(y) & (y != -9)]).max()
    if form == "mfma":
        # both runs sum the same fixed-point phenotype values over the same individuals
        tol_s, tol_f = (mapped + 1) * scale * 2.0 ** -22, TOL_F
    else:
        # each run is within the FP32 bound of the exact sums, and each F within TOL_F of the oracle
        tol_s, tol_f = 2 * (1e-7 * nid * scale + 1e-9), 2 * TOL_F
    assert (np.abs(s1[sel] - msum) <= tol_s).all()
    ok = (c0[sel] > 0).sum(1) >= 2
    assert rel_err(F1[sel][ok], F0[sel][ok]).max() <= tol_f
    assert rel_err(G1[sel][ok], G0[sel][ok]).max() <= tol_f
