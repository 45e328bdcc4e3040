"""GPU tests of the batched persistent launch (episcan.h epi_batch_*): many scans of uneven N and M,
intra- and inter-chromosome blocks and several phenotypes per dataset in one captured launch. Every
scan's hits must equal those of a single-scan epi_scan on the same data, and agree with the FP64
oracle."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL_F = 1e-5
T1, T2 = 3.6, 4.6

# (nid, chromosome sizes, missing-call rate, phenotype distributions)
SPECS = [
    (300, [180], 0.0, [0]),
    (517, [150, 130], 0.01, [0, 2]),
    (1000, [120, 90, 70], 0.0, [0, 3, 4]),
    (259, [260], 0.005, [1]),
    (777, [200, 160], 0.0, [0]),
    (1500, [330], 0.01, [2, 0]),
]


def _data():
    import epigpu_amd as E
    out = []
    for i, (nid, sizes, gmiss, dists) in enumerate(SPECS):
        nsnp = sum(sizes)
        words = E.synth_genotypes(0xBA7C0 + i, nid, nsnp, 0.05, 0.5, gmiss)
        phen = np.stack([E.synth_phenotype(0xBA7C0 + i, r, nid, d, 0.03) for r, d in enumerate(dists)])
        starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
        out.append((nid, words, phen, starts, np.asarray(sizes, dtype=np.int32)))
    return out


@pytest.fixture(scope="module")
def batch_data():
    return _data()


def _make_batch(scanner, data):
    import epigpu_amd as E
    b = E.Batch(scanner)
    scans = []
    for nid, words, phen, starts, sizes in data:
        d = b.add_dataset(words, nid, starts, sizes)
        for r, s in enumerate(b.add_phenotypes(d, phen)):
            scans.append((len(scans), d, r))
            assert s == scans[-1][0]
    return b, scans


@pytest.mark.parametrize("form,rows", [("plain", 32), ("plain", 16), ("mfma", 32), ("mfma", 64)])
def test_batch_matches_single_scans(scanner, batch_data, form, rows):
    """Both tile heights of each form (MFMA 64: 512-thread workgroups, 64 x 32 GENERAL tiles)."""
    b, scans = _make_batch(scanner, batch_data)
    b.set_schedule(form, fast_rows=rows)
    print(f"batch form={form} rows={rows} scans={len(scans)}", flush=True)
    st = b.scan(test="i", form=form, t1=T1, t2=T2)
    assert not st["overflow"]
    total = 0
    exp_pairs = 0
    for s, d, r in scans:
        nid, words, phen, starts, sizes = batch_data[d]
        g = b.hits(s)
        total += g["nhits"]
        scanner.set_genotypes(words, nid)
        scanner.set_phenotypes(phen)
        scanner.set_chromosomes(starts, sizes)
        one = scanner.scan(test="i", form=form, t1=T1, t2=T2, phen=r)
        exp_pairs += one["pairs"]
        ka = list(zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist()))
        kb = list(zip(one["hits"]["snp1"].tolist(), one["hits"]["snp2"].tolist()))
        assert ka == kb, f"scan {s}: batch and single-scan hit lists differ"
        assert (g["block"] == one["block"]).all()
        assert (g["hits"]["df1"] == one["hits"]["df1"]).all() and (g["hits"]["df2"] == one["hits"]["df2"]).all()
        assert np.abs(g["F64"] - one["F64"]).max(initial=0) <= 1e-12 * max(1.0, np.abs(one["F64"]).max(initial=0))
        assert np.abs(g["F2_64"] - one["F2_64"]).max(initial=0) <= 1e-12 * max(1.0,
                                                                               np.abs(one["F2_64"]).max(initial=0))
        assert g["nhits"] > 0
    assert total == st["nhits"]
    assert st["pairs"] == exp_pairs
    b.close()


def test_batch_against_oracle(scanner, batch_data):
    """Two scans of the MFMA batch (ragged nid, missing calls, 3 blocks) against the FP64 oracle."""
    b, scans = _make_batch(scanner, batch_data)
    b.scan(test="i", form="mfma", t1=T1, t2=T2)
    for s, d, r in [sc for sc in scans if sc[1] in (1, 2)][:3]:
        nid, words, phen, starts, sizes = batch_data[d]
        g = b.hits(s)
        n_ref, ref = O.scan(O.unpack(words, nid), phen[r], starts, sizes, "i", T1, T2)
        gs = {(int(a), int(c)): (f, f2) for a, c, f, f2 in zip(g["hits"]["snp1"], g["hits"]["snp2"], g["F64"],
                                                               g["F2_64"])}
        rs = {(int(a), int(c)): (f, f2) for a, c, f, f2 in zip(ref["snp1"], ref["snp2"], ref["F"], ref["F2"])}
        for pr in set(gs) ^ set(rs):  # only near-threshold ties may differ
            f, f2 = rs.get(pr, gs.get(pr))
            assert abs(f - T1) <= 1e-5 * T1 or abs(f2 - T2) <= 1e-5 * T2, f"scan {s}: non-tie difference {pr}"
        common = sorted(set(gs) & set(rs))
        assert len(common) > 20
        assert max(rel_err(gs[p][0], rs[p][0]) for p in common) <= TOL_F
        assert max(rel_err(gs[p][1], rs[p][1]) for p in common) <= TOL_F
    b.close()


def test_batch_graph_replay_and_recapture(scanner, batch_data):
    """Replays of the captured graph give the same hits; new settings capture a new graph."""
    b, scans = _make_batch(scanner, batch_data[:3])
    first = b.scan(test="i", form="mfma", t1=T1, t2=T2)
    h0 = [b.hits(s) for s, _, _ in scans]
    for _ in range(3):
        again = b.scan(test="i", form="mfma", t1=T1, t2=T2)
        assert again["captures"] == 1 and again["nhits"] == first["nhits"]
        for (s, _, _), h in zip(scans, h0):
            assert (b.hits(s)["hits"] == h["hits"]).all()
    f_only = b.scan(test="f", form="mfma", t1=T1, t2=T2)
    assert f_only["captures"] == 2 and f_only["nhits"] > 0
    plain = b.scan(test="i", form="plain", t1=T1, t2=T2)
    assert plain["captures"] == 3
    back = b.scan(test="i", form="mfma", t1=T1, t2=T2)
    assert back["captures"] == 4 and back["nhits"] == first["nhits"]
    b.close()


def test_batch_overflow_and_errors(scanner, batch_data):
    import epigpu_amd as E
    b, scans = _make_batch(scanner, batch_data[:2])
    full = b.scan(test="f", form="mfma", t1=1.5)
    b.set_hit_capacity(50)
    small = b.scan(test="f", form="mfma", t1=1.5)
    assert full["nhits"] > 1000 and small["nhits"] == full["nhits"] and small["overflow"]
    assert sum(b.hit_count(s)[0] for s, _, _ in scans) == 50
    with pytest.raises(E.EpiError):
        b.add_phenotypes(0, batch_data[0][2])  # phenotypes already added
    with pytest.raises(E.EpiError):
        b.hits(len(scans))  # no such scan
    empty = E.Batch(scanner)
    with pytest.raises(E.EpiError):
        empty.scan()
    empty.close()
    b.close()


SCHEDULES = {
    "mfma": [dict(fast_rows=64), dict(fast_rows=64, super_a=1, super_b=1), dict(fast_rows=32, super_a=16, super_b=2),
             dict(fast_rows=64, dense_missing=1), dict(fast_rows=32, dense_missing=1),
             dict(fast_rows=64, dense_missing=2), dict(fast_rows=32, dense_missing=2)],
    "plain": [dict(fast_rows=16), dict(fast_rows=32, super_a=4, super_b=8)],
}


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_single_scan_schedules_agree(scanner, batch_data, form):
    """Every tile shape and issue order gives the default schedule's hits (missing calls, 2 blocks)."""
    nid, words, phen, starts, sizes = batch_data[1]
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(phen)
    scanner.set_chromosomes(starts, sizes)
    default = scanner.schedule(form)
    ref = scanner.scan(test="i", form=form, t1=T1, t2=T2, phen=1)
    try:
        for kw in SCHEDULES[form]:
            scanner.set_schedule(form, **kw)
            print(f"schedule {form} {kw}", flush=True)
            g = scanner.scan(test="i", form=form, t1=T1, t2=T2, phen=1)
            assert len(g["hits"]) == len(ref["hits"]) and g["pairs"] == ref["pairs"]
            assert (g["hits"]["snp1"] == ref["hits"]["snp1"]).all() and (g["hits"]["snp2"] == ref["hits"]["snp2"]).all()
            assert np.abs(g["F64"] - ref["F64"]).max() <= 1e-12 * ref["F64"].max()
            assert np.abs(g["F2_64"] - ref["F2_64"]).max() <= 1e-12 * ref["F2_64"].max()
    finally:
        scanner.set_schedule(form, **default)


def test_batch_autotune(scanner, batch_data):
    """The autotuner measures candidates, keeps the fastest and the batch then scans with it."""
    b, scans = _make_batch(scanner, batch_data[:4])
    base = b.scan(test="i", form="mfma", t1=T1, t2=T2)
    h0 = [b.hits(s) for s, _, _ in scans]
    res = b.autotune(test="i", form="mfma", t1=T1, t2=T2, reps=1)
    assert len(res["log"]) >= 8
    assert res["best_ms"] == min(e["ms"] for e in res["log"])
    assert b.schedule("mfma") == res["best"]
    after = b.scan(test="i", form="mfma", t1=T1, t2=T2)
    assert after["nhits"] == base["nhits"]
    for (s, _, _), h in zip(scans, h0):
        g = b.hits(s)
        assert (g["hits"]["snp1"] == h["hits"]["snp1"]).all() and (g["hits"]["snp2"] == h["hits"]["snp2"]).all()
        assert np.abs(g["F64"] - h["F64"]).max(initial=0) <= 1e-12 * max(1.0, np.abs(h["F64"]).max(initial=0))
    # every schedule the tuner can pick scans the same pairs
    for sched in res["log"]:
        b.set_schedule("mfma", **{k: sched[k] for k in ("fast_rows", "super_a", "super_b", "queues", "order",
                                                        "dense_missing")})
        st = b.scan(test="i", form="mfma", t1=T1, t2=T2)
        assert st["nhits"] == base["nhits"] and st["pairs"] == base["pairs"]
    b.close()


def test_single_scan_autotune(scanner, batch_data):
    """The single-scan autotuner keeps the fastest measured schedule and the hits do not change."""
    nid, words, phen, starts, sizes = batch_data[2]
    scanner.set_genotypes(words, nid)
    scanner.set_phenotypes(phen)
    scanner.set_chromosomes(starts, sizes)
    default = scanner.schedule("mfma")
    ref = scanner.scan(test="i", form="mfma", t1=T1, t2=T2)
    try:
        res = scanner.autotune(test="i", form="mfma", t1=T1, t2=T2, reps=2)
        assert len(res["log"]) >= 5
        assert res["best_ms"] == min(e["ms"] for e in res["log"])
        assert scanner.schedule("mfma") == res["best"]
        g = scanner.scan(test="i", form="mfma", t1=T1, t2=T2)
        assert (g["hits"]["snp1"] == ref["hits"]["snp1"]).all() and (g["hits"]["snp2"] == ref["hits"]["snp2"]).all()
        assert np.abs(g["F64"] - ref["F64"]).max(initial=0) <= 1e-12 * max(1.0, np.abs(ref["F64"]).max(initial=0))
    finally:
        scanner.set_schedule("mfma", **default)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_batch_safe_mode_matches_single_scans(scanner, batch_data, form):
    """Safe mode in the batch (-9 a phenotype value, test 'i' without nfac > 6) equals safe single scans."""
    import epigpu_amd as E
    b = E.Batch(scanner)
    data = []
    for d, (nid, words, phen, starts, sizes) in enumerate(batch_data[:3]):
        ph = phen.copy()
        ph[:, ::7] = -9.0  # -9 is a value here
        data.append((nid, words, ph, starts, sizes))
        di = b.add_dataset(words, nid, starts, sizes)
        b.add_phenotypes(di, ph, nine_is_value=True)
    st = b.scan(test="i", form=form, t1=T1, t2=T2, safe=True)
    assert not st["overflow"]
    s = 0
    for nid, words, ph, starts, sizes in data:
        scanner.set_genotypes(words, nid)
        scanner.set_phenotypes(ph, nine_is_value=True)
        scanner.set_chromosomes(starts, sizes)
        for r in range(len(ph)):
            g = b.hits(s)
            one = scanner.scan(test="i", form=form, t1=T1, t2=T2, phen=r, safe=True)
            assert len(g["hits"]) == len(one["hits"]), f"scan {s}: {len(g['hits'])} != {len(one['hits'])}"
            assert (g["hits"]["snp1"] == one["hits"]["snp1"]).all() and (g["hits"]["snp2"] == one["hits"]["snp2"]).all()
            assert np.abs(g["F64"] - one["F64"]).max(initial=0) <= 1e-12 * max(1.0, np.abs(one["F64"]).max(initial=0))
            s += 1
    # and the safe batch differs from the normal one on this data (-9 counts as a value)
    normal = E.Batch(scanner)
    for nid, words, ph, starts, sizes in data:
        normal.add_phenotypes(normal.add_dataset(words, nid, starts, sizes), ph)
    assert normal.scan(test="i", form=form, t1=T1, t2=T2)["nhits"] != st["nhits"]
    normal.close()
    b.close()
