"""GPU parity of phenotypes whose values span a wide range within one vector: single outliers 30 to 10,000
SDs out, tens of them at once, and raw skewed traits, on both kernel forms and both MFMA tile kinds, against
the FP64 oracle through the per-pair dump and through the production hit list. The MFMA form carries up to
32 far values exactly beside its fixed-point encoding; a phenotype with more far values than that is refused
by MFMA scans (EPI_ERR_RANGE) and scanned by the plain form. Also: a failed phenotype upload leaves no
phenotypes behind.

Bar: counts bit-exact; F and F2 within 1e-5 of the oracle (rel_err); hit sets equal apart from listed
threshold ties."""
import numpy as np
import pytest

import oracle as O
from conftest import hits_vs_oracle, make_dataset, rel_err

pytestmark = pytest.mark.gpu

TOL_F = 1e-5
EPI_ERR_STATE, EPI_ERR_RANGE = -4, -6


def with_outliers(seed, nid, nsnp, sds, drop_third, gmiss=0.0, dist=0):
    """A dataset whose phenotype has values sds[i] SDs out (alternating signs) at random individuals.
    drop_third: each of those individuals misses its call on about a third of the SNPs, so that pairs
    both include and exclude it."""
    import epigpu_amd as E
    words, y = make_dataset(seed, nid, nsnp, maf=(0.05, 0.5), gmiss=gmiss, dist=dist, pmiss=0.01)
    rng = np.random.default_rng(seed)
    pres = np.flatnonzero(np.isfinite(y))
    idx = rng.choice(pres, len(sds), replace=False)
    m, s = y[pres].astype(np.float64).mean(), y[pres].astype(np.float64).std()
    sign = np.where(np.arange(len(sds)) % 2 == 0, 1.0, -1.0)
    y[idx] = (m + sign * np.asarray(sds, dtype=np.float64) * s).astype(np.float32)
    if drop_third:
        codes = O.unpack(words, nid)
        for k in idx:
            codes[rng.random(nsnp) < 1.0 / 3.0, k] = 3
        words = E.pack(codes)
    return words, y


def dump_check(scanner, words, y, nid, form):
    nsnp = words.shape[0]
    counts, sums, F, F2 = scanner.pair_dump(0, nsnp, 0, nsnp, test="i", form=form)
    aa, bb = np.meshgrid(np.arange(nsnp), np.arange(nsnp), indexing="ij")
    sel = aa > bb
    codes = O.unpack(words, nid)
    ref = O.pairs(codes, y, np.stack([aa[sel], bb[sel]], axis=1))
    assert (counts[sel] == ref["counts"]).all(), "counts differ"
    ok = ref["df1"] >= 1
    ef = rel_err(F[sel][ok], ref["F"][ok])
    ef2 = rel_err(F2[sel][ok], ref["F2"][ok])
    assert ef.max() <= TOL_F, f"max F error {ef.max():.3g}"
    assert ef2.max() <= TOL_F, f"max F2 error {ef2.max():.3g}"
    return ref, ok, ef.max(), ef2.max()


def scan_check(scanner, words, y, nid, form, ref, ok):
    """The production hit list at thresholds that about 2 % of the pairs pass, against the oracle's scan."""
    nsnp = words.shape[0]
    t1 = float(np.float32(np.quantile(ref["F"][ok], 0.99)))
    t2 = float(np.float32(np.quantile(ref["F2"][ok], 0.99)))
    scanner.set_chromosomes([0], [nsnp])
    g = scanner.scan(test="i", form=form, t1=t1, t2=t2)
    n, ties = hits_vs_oracle(g, O.unpack(words, nid), y, [0], [nsnp], "i", t1, t2)
    assert n > 0
    return n, ties


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("kind", ["fast", "general"])
@pytest.mark.parametrize("sd", [30, 100, 1e3, 1e4])
def test_one_outlier(scanner, form, kind, sd):
    """One value sd SDs out. FAST tiles: no missing calls, the outlier is in every pair. GENERAL tiles: it
    misses a call on a third of the SNPs, so a third of the pairs drop it and their F rests on the rest."""
    nid, nsnp = 3000, 160
    words, y = with_outliers(0xA17 + int(sd), nid, nsnp, [sd], drop_third=(kind == "general"))
    print(f"one outlier sd={sd:g} kind={kind} form={form}", flush=True)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    ref, ok, ef, ef2 = dump_check(scanner, words, y, nid, form)
    scan_check(scanner, words, y, nid, form, ref, ok)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("rows", [32, 64])
def test_many_outliers(scanner, form, rows):
    """20 far values (both outlier code words of a SNP in use) from 40 to 1e5 SDs, half of them missing calls
    on a third of the SNPs, with GENERAL and FAST tiles in one scan (genotype calls missing in only some SNP
    ranges), at both MFMA tile heights."""
    import epigpu_amd as E
    nid, nsnp = 2500, 200
    sds = np.geomspace(40, 1e5, 20)
    words, y = with_outliers(0xB0B, nid, nsnp, sds, drop_third=False)
    codes = O.unpack(words, nid)
    rng = np.random.default_rng(7)
    far = np.argsort(-np.abs(y - np.nanmedian(y)))[:10]
    snps = np.arange(100, 200)  # calls missing only in the upper half: tiles of both kinds
    for k in far:
        codes[snps[rng.random(len(snps)) < 1.0 / 3.0], k] = 3
    words = E.pack(codes)
    print(f"many outliers form={form} rows={rows}", flush=True)
    scanner.set_schedule(form, fast_rows=rows if form == "mfma" else (32 if rows == 64 else 16))
    try:
        scanner.set_genotypes(words, nid)
        scanner.set_phenotypes(y)
        ref, ok, ef, ef2 = dump_check(scanner, words, y, nid, form)
        scan_check(scanner, words, y, nid, form, ref, ok)
    finally:
        scanner.set_schedule(form, fast_rows=32)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("trait", ["lognormal1", "lognormal2", "gamma", "t3", "binary0.002"])
def test_skewed_raw_traits(scanner, form, trait):
    """Raw (untransformed) skewed and heavy-tailed traits and a rare binary trait, with 1 % missing calls."""
    nid, nsnp = 4000, 128
    words, _ = make_dataset(0xC0DE, nid, nsnp, gmiss=0.01)
    rng = np.random.default_rng(["lognormal1", "lognormal2", "gamma", "t3", "binary0.002"].index(trait) + 90)
    y = {"lognormal1": lambda: np.exp(rng.standard_normal(nid)),
         "lognormal2": lambda: np.exp(2.0 * rng.standard_normal(nid)),
         "gamma": lambda: rng.gamma(2.0, 1.0, nid),
         "t3": lambda: rng.standard_t(3, nid),
         "binary0.002": lambda: (rng.random(nid) < 0.002).astype(np.float64)}[trait]().astype(np.float32)
    y[rng.random(nid) < 0.02] = np.nan
    print(f"raw trait {trait} form={form}", flush=True)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    ref, ok, ef, ef2 = dump_check(scanner, words, y, nid, form)
    if trait != "binary0.002":
        scan_check(scanner, words, y, nid, form, ref, ok)


def test_mfma_refuses_beyond_range(scanner):
    """40 values 1e4 SDs out: more far values than the MFMA form carries exactly. MFMA scans, dumps and batch
    scans of that phenotype fail with EPI_ERR_RANGE; the plain form scans it and matches the oracle; a
    second, ordinary phenotype of the same upload still scans in the MFMA form."""
    import epigpu_amd as E
    nid, nsnp = 1500, 96
    words, y = with_outliers(0xD00D, nid, nsnp, [1e4] * 40, drop_third=True)
    y2 = E.synth_phenotype(5, 1, nid, 0, 0.01)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(np.stack([y, y2]))
    scanner.set_chromosomes([0], [nsnp])
    for call in (lambda: scanner.scan(test="i", form="mfma", phen=0),
                 lambda: scanner.pair_dump(0, nsnp, 0, nsnp, form="mfma", phen=0)):
        with pytest.raises(E.EpiError) as ei:
            call()
        assert ei.value.code == EPI_ERR_RANGE and "plain form" in str(ei.value)
    g = scanner.scan(test="i", form="mfma", phen=1, t1=3.0, t2=4.0)
    hits_vs_oracle(g, O.unpack(words, nid), y2, [0], [nsnp], "i", 3.0, 4.0)
    scanner.set_phenotypes(y)
    ref, ok, _, _ = dump_check(scanner, words, y, nid, "plain")
    scan_check(scanner, words, y, nid, "plain", ref, ok)
    with E.Batch(scanner) as b:
        d = b.add_dataset(words, nid)
        b.add_phenotypes(d, y)
        with pytest.raises(E.EpiError) as ei:
            b.scan(test="i", form="mfma")
        assert ei.value.code == EPI_ERR_RANGE
        b.scan(test="i", form="plain", t1=3.0, t2=4.0)
        assert b.hit_count(0)[0] > 0


def test_batch_with_outliers_equals_single_scans(scanner):
    """Outlier phenotypes in the persistent batch launch (two phenotypes of one dataset) give the single
    scans' hits and FP64 values."""
    import epigpu_amd as E
    nid, nsnp = 2000, 180
    words, y1 = with_outliers(0xE1, nid, nsnp, [1e4, 300, 50], drop_third=True)
    _, y2 = with_outliers(0xE2, nid, nsnp, list(np.geomspace(40, 1e4, 18)), drop_third=False)
    t1, t2 = 3.0, 4.5
    single = []
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(np.stack([y1, y2]))
    scanner.set_chromosomes([0, 90], [90, 90])
    for r in range(2):
        single.append(scanner.scan(test="i", form="mfma", t1=t1, t2=t2, phen=r))
    with E.Batch(scanner) as b:
        d = b.add_dataset(words, nid, [0, 90], [90, 90])
        s = b.add_phenotypes(d, np.stack([y1, y2]))
        b.scan(test="i", form="mfma", t1=t1, t2=t2)
        for r in range(2):
            g = b.hits(s[r])
            assert g["nhits"] == single[r]["nhits"] > 0
            assert (g["hits"] == single[r]["hits"]).all()
            assert np.array_equal(g["F64"], single[r]["F64"]) and np.array_equal(g["F2_64"], single[r]["F2_64"])
            hits_vs_oracle(g, O.unpack(words, nid), [y1, y2][r], [0, 90], [90, 90], "i", t1, t2)


@pytest.mark.parametrize("bad", ["inf", "all_missing"])
def test_failed_phenotype_upload_leaves_none(scanner, bad):
    """A phenotype upload that fails (an infinite value, or fewer than two present values) drops the old
    phenotypes too: the next scan reports EPI_ERR_STATE instead of scanning stale or partial state, and a
    good upload afterwards scans normally. In a batch the failed dataset can take phenotypes again."""
    import epigpu_amd as E
    nid, nsnp = 500, 64
    words, y = make_dataset(0xF00, nid, nsnp, gmiss=0.01)
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes([0], [nsnp])
    scanner.scan(test="i", form="mfma")
    yb = y.copy()
    if bad == "inf":
        yb[7] = np.inf
    else:
        yb[:] = -9.0
    with pytest.raises(E.EpiError) as ei:
        scanner.set_phenotypes(np.stack([y, yb]))
    assert ei.value.code == EPI_ERR_RANGE
    for form in ("mfma", "plain"):
        with pytest.raises(E.EpiError) as ei:
            scanner.scan(test="i", form=form)
        assert ei.value.code == EPI_ERR_STATE
    scanner.set_phenotypes(y)
    g = scanner.scan(test="i", form="mfma", t1=3.0, t2=4.0)
    hits_vs_oracle(g, O.unpack(words, nid), y, [0], [nsnp], "i", 3.0, 4.0)
    with E.Batch(scanner) as b:
        d = b.add_dataset(words, nid)
        with pytest.raises(E.EpiError):
            b.add_phenotypes(d, yb)
        s = b.add_phenotypes(d, y)
        b.scan(test="i", form="mfma", t1=3.0, t2=4.0)
        hits_vs_oracle(b.hits(s[0]), O.unpack(words, nid), y, [0], [nsnp], "i", 3.0, 4.0)
