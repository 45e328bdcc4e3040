"""GPU parity of the production hit path at the benchmark workloads' full sizes.

Workloads (epigpu_amd.workloads): c2 N=4,000, M=20,000 over 2 chromosomes, 1 % missing calls, 2 % missing
phenotypes (GENERAL tiles); c3 N=8,000, M=50,000, standardized Gamma(2,1) phenotype with 10 planted pairs (FAST
tiles); c4 64 scans with N in [300, 6000] and M in [1000, 30000] over 1-4 chromosomes in one persistent launch;
c5 N=20,000, M=200,000, 1 % missing calls, 4 phenotypes in one launch.

For each: (i) a block of the per-pair dump against the oracle (counts exact, F and F2 within 1e-5);
(ii) every hit the scan reports (the FP32 screen, the FP64 decision and the compaction, not the dump path)
has the oracle's df1, df2 and F, F2 within 1e-5, and passes the oracle's hit rule unless it is a threshold
tie; (iii) in windows of 256 x 256 SNPs, the oracle's hits equal the reported ones apart from listed ties.
c2, c3 and c4 scan at thresholds near p < 1e-4, where the null alone gives hundreds of thousands of hits."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5
WIN = 256


def sub_pairs(words, nid, y, pairs):
    """Oracle statistics for (a, b) rows of pairs, unpacking only the SNPs involved."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    snps = np.unique(pairs)
    codes = O.unpack(np.ascontiguousarray(words[snps]), nid)
    return O.pairs(codes, y, np.searchsorted(snps, pairs).astype(np.int32))


def oracle_hit(o, test, t1, t2):
    ok = (o["df1"] >= 1) & (o["df2"] >= 1)
    if test == "f":
        return ok & (o["F"] > t1)
    return ok & ((o["F"] > t1) | (o["F2"] > t2)) & (o["nfac"] > 6)


def near(o, test, t1, t2):
    x = np.abs(o["F"] - t1) <= TOL * max(t1, 1.0)
    if test == "i":
        x |= np.abs(o["F2"] - t2) <= TOL * max(t2, 1.0)
    return x


def check_reported(g, words, nid, y, test, t1, t2):
    """(ii): every reported hit against the oracle. Returns the number of ties among them."""
    pr = np.stack([g["hits"]["snp1"], g["hits"]["snp2"]], axis=1)
    assert len(pr) > 0
    assert len({(a, b) for a, b in pr.tolist()}) == len(pr), "a pair is reported twice"
    assert (pr[:, 0] > pr[:, 1]).all()
    o = sub_pairs(words, nid, y, pr)
    assert (g["hits"]["df1"] == o["df1"]).all() and (g["hits"]["df2"] == o["df2"]).all()
    ef = rel_err(g["F64"], o["F"])
    assert ef.max() <= TOL, f"reported F off the oracle's by {ef.max():.3g}"
    if test == "i":
        ef2 = rel_err(g["F2_64"], o["F2"])
        assert ef2.max() <= TOL, f"reported F2 off the oracle's by {ef2.max():.3g}"
    bad = ~oracle_hit(o, test, t1, t2)
    assert (near(o, test, t1, t2) | ~bad).all(), "a reported pair is no hit for the oracle"
    return int(bad.sum())


def check_windows(g, words, nid, y, chr_starts, chr_sizes, test, t1, t2, windows):
    """(iii): all pairs of each window (a in [a0, a0+WIN), b in [b0, b0+WIN), a > b) through the oracle; its
    hits must be the reported hits of the window apart from ties. Returns (oracle hits, ties)."""
    rep = {(a, b) for a, b in zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist())}
    tot = ties = 0
    for a0, b0 in windows:
        aa, bb = np.meshgrid(np.arange(a0, a0 + WIN), np.arange(b0, b0 + WIN), indexing="ij")
        sel = aa > bb
        pr = np.stack([aa[sel], bb[sel]], axis=1)
        o = sub_pairs(words, nid, y, pr)
        want = oracle_hit(o, test, t1, t2)
        got = np.array([(a, b) in rep for a, b in pr.tolist()])
        diff = want != got
        assert (~diff | near(o, test, t1, t2)).all(), \
            f"window ({a0}, {b0}): {int(diff.sum())} pairs differ, first {pr[diff][:5].tolist()}"
        tot += int(want.sum())
        ties += int(diff.sum())
    return tot, ties


def dump_block(scanner, words, nid, y, a0, b0, n, form, phen=0):
    """(i): the n x n dump block at (a0, b0) against the oracle."""
    counts, sums, F, F2 = scanner.pair_dump(a0, a0 + n, b0, b0 + n, test="i", form=form, phen=phen)
    aa, bb = np.meshgrid(np.arange(a0, a0 + n), np.arange(b0, b0 + n), indexing="ij")
    sel = aa > bb
    o = sub_pairs(words, nid, y, np.stack([aa[sel], bb[sel]], axis=1))
    assert (counts[sel] == o["counts"]).all()
    ok = o["df1"] >= 1
    assert rel_err(F[sel][ok], o["F"][ok]).max() <= TOL
    assert rel_err(F2[sel][ok], o["F2"][ok]).max() <= TOL


def p4(nid):
    """F thresholds at p < 1e-4 of the 8df and 4df tests (as FP32, the scan's threshold type)."""
    import epigpu_amd as E
    return float(np.float32(E.f_isf(1e-4, 8, nid - 9))), float(np.float32(E.f_isf(1e-4, 4, nid - 9)))


def windows_of(nsnp, chr_starts, k, seed):
    """k window corners: one on the first diagonal block, the others random (a0 >= b0)."""
    rng = np.random.default_rng(seed)
    out = [(chr_starts[0] + 64, chr_starts[0] + 64)]
    while len(out) < k:
        a0 = int(rng.integers(0, nsnp - WIN))
        b0 = int(rng.integers(0, a0 + 1))
        out.append((a0, b0))
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,form", [("c2", "plain"), ("c2", "mfma"), ("c3", "mfma"), ("c3", "plain")])
def test_full_size_workload(scanner, name, form):
    from epigpu_amd import workloads as W
    w = W.get(name).generate()
    y = w.y[0]
    t1, t2 = p4(w.nid)
    print(f"{name} form={form}: N={w.nid} M={w.nsnp} t1={t1} t2={t2}", flush=True)
    scanner.set_genotypes(w.words, w.nid)
    scanner.set_phenotypes(y)
    scanner.set_chromosomes(w.chr_starts, w.chr_sizes)
    scanner.set_hit_capacity(1 << 23)
    try:
        g = scanner.scan(test="i", form=form, t1=t1, t2=t2)
        assert not g["overflow"] and g["pairs"] == w.pairs
        print(f"{name} {form}: {g['nhits']} hits in {g['kernel_ms']:.1f} ms", flush=True)
        ties = check_reported(g, w.words, w.nid, y, "i", t1, t2)
        nh, wt = check_windows(g, w.words, w.nid, y, w.chr_starts, w.chr_sizes, "i", t1, t2,
                               windows_of(w.nsnp, w.chr_starts, 4, 11))
        print(f"{name} {form}: reported hits checked ({ties} ties), windows {nh} oracle hits ({wt} ties)", flush=True)
        assert nh > 20
        dump_block(scanner, w.words, w.nid, y, w.nsnp - 64, w.nsnp - 200, 64, form)
    finally:
        scanner.set_hit_capacity(1 << 22)


@pytest.mark.timeout(900)
def test_batch_workload_c4(scanner):
    """The 64-scan batch in one captured persistent launch (MFMA form); 4 scans (the smallest and largest
    cohorts and two others) checked against the oracle: every reported hit and 2 windows each."""
    import epigpu_amd as E
    from epigpu_amd import workloads as W
    bw = W.get_batch("c4").generate()
    t1, t2 = 3.99, 5.9  # about p < 1e-4 for every cohort size of the batch
    with E.Batch(scanner) as b:
        ids = []
        for w in bw.scans:
            d = b.add_dataset(w.words, w.nid, w.chr_starts, w.chr_sizes)
            ids.append(b.add_phenotypes(d, w.y[0])[0])
        b.set_hit_capacity(1 << 23)
        st = b.scan(test="i", form="mfma", t1=t1, t2=t2)
        assert not st["overflow"] and st["pairs"] == bw.pairs
        print(f"c4: {st['nhits']} hits in {st['kernel_ms']:.1f} ms", flush=True)
        nids = [w.nid for w in bw.scans]
        pick = sorted({int(np.argmin(nids)), int(np.argmax(nids)), 7, 40})
        for i in pick:
            w = bw.scans[i]
            g = b.hits(ids[i])
            ties = check_reported(g, w.words, w.nid, w.y[0], "i", t1, t2)
            nh, wt = check_windows(g, w.words, w.nid, w.y[0], w.chr_starts, w.chr_sizes, "i", t1, t2,
                                   windows_of(w.nsnp, w.chr_starts, 2, i))
            print(f"c4 scan {i} (N={w.nid}, M={w.nsnp}): {g['nhits']} hits ({ties} ties), windows {nh} ({wt} ties)",
                  flush=True)


@pytest.mark.timeout(1200)
def test_multi_phenotype_workload_c5(scanner):
    """c5's four phenotypes (N(0,1), raw Gamma(2,1), t(5), 0/1 at prevalence 0.2) as one batched launch at the
    workload's thresholds (6.5 / 10.5): every reported hit of every phenotype against the oracle, a 64 x 64 dump
    block of each, and 2 windows of each (at these thresholds a window rarely holds a hit; the window check at
    p < 1e-4 is the c2, c3 and c4 tests')."""
    import epigpu_amd as E
    from epigpu_amd import workloads as W
    w = W.get("c5").generate()
    print(f"c5: N={w.nid} M={w.nsnp}, 4 phenotypes", flush=True)
    with E.Batch(scanner) as b:
        d = b.add_dataset(w.words, w.nid, w.chr_starts, w.chr_sizes)
        s = b.add_phenotypes(d, w.y)
        st = b.scan(test="i", form="mfma", t1=w.t1, t2=w.t2)
        assert not st["overflow"] and st["pairs"] == 4 * w.pairs
        print(f"c5: {st['nhits']} hits in {st['kernel_ms']:.0f} ms", flush=True)
        for r in range(4):
            g = b.hits(s[r])
            ties = check_reported(g, w.words, w.nid, w.y[r], "i", w.t1, w.t2)
            nh, wt = check_windows(g, w.words, w.nid, w.y[r], w.chr_starts, w.chr_sizes, "i", w.t1, w.t2,
                                   windows_of(w.nsnp, w.chr_starts, 2, 50 + r))
            print(f"c5 phenotype {r}: {g['nhits']} hits ({ties} ties), windows {nh} ({wt} ties)", flush=True)
    scanner.set_genotypes(w.words, w.nid)
    scanner.set_phenotypes(w.y)
    for r in range(4):
        dump_block(scanner, w.words, w.nid, w.y[r], 150000, 90000 + 64 * r, 64, "mfma", phen=r)
