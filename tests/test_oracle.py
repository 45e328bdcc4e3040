"""Pins the FP64 oracle (CPU): golden vectors from library routines, closed forms,
brute force and the invariances of the statistics. No GPU."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from conftest import make_dataset, rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def one_pair(ga, gb, y):
    codes = np.stack([np.asarray(ga, np.int8), np.asarray(gb, np.int8)])
    return O.pairs(codes, np.asarray(y, np.float32), np.array([[0, 1]], np.int32))


def test_oracle_F_matches_f_oneway():
    g = np.load(os.path.join(GOLD, "anova.npz"))
    assert len(g["F"]) >= 30
    for ga, gb, y, F, nf in zip(g["ga"], g["gb"], g["y"], g["F"], g["nfac"]):
        r = one_pair(ga, gb, y)
        assert r["nfac"][0] == nf
        assert r["df1"][0] == nf - 1 and r["df2"][0] == len(y) - nf
        assert abs(r["F"][0] - F) <= 1e-9 * max(1.0, abs(F))


def test_oracle_F2_matches_two_way_decomposition():
    """Proportional cell counts: reference SSI == SS_cells - SS_A - SS_B."""
    g = np.load(os.path.join(GOLD, "interaction.npz"))
    for ga, gb, y, F2, F in zip(g["ga"], g["gb"], g["y"], g["F2"], g["F"]):
        r = one_pair(ga, gb, y)
        assert abs(r["F2"][0] - F2) <= 1e-9 * max(1.0, abs(F2))
        assert abs(r["F"][0] - F) <= 1e-9 * max(1.0, abs(F))


def test_interaction_ss_bounded_by_between_cells_ss():
    """SSI <= SSB for every table, balanced or not: SSI is |e - Pr e - Pc e|^2 for the cell deviations e
    and its n-weighted projections on row- and column-constant vectors, which is |e|^2 - |Pr e|^2 -
    |Pc e|^2 + 2<Pr e, Pc e> <= |e|^2 = SSB. Hence F2 <= df1 * F / 4, the bound the GPU screen uses to
    skip the F2 terms. Checked on unbalanced, heavy-tailed and missing-data tables, with equality
    approached when the deviations are pure interaction."""
    rng = np.random.default_rng(11)
    worst = 0.0
    for trial in range(30):
        nid = int(rng.integers(30, 300))
        maf = rng.uniform(0.02, 0.5, 24)
        codes = ((rng.random((24, nid)) < maf[:, None]).astype(np.int8)
                 + (rng.random((24, nid)) < maf[:, None]).astype(np.int8))
        codes[rng.random((24, nid)) < 0.05] = 3
        y = (rng.standard_t(2, nid) if trial % 2 else rng.standard_normal(nid)).astype(np.float32)
        if trial % 3 == 0:  # pure interaction: y follows (g_a - 1)(g_b - 1)
            y = y * 0.01 + ((codes[0].astype(np.float32) - 1) * (codes[12].astype(np.float32) - 1))
        aa, bb = np.meshgrid(np.arange(12), np.arange(12, 24), indexing="ij")
        r = O.pairs(codes, y, np.stack([aa.ravel(), bb.ravel()], 1).astype(np.int32))
        ok = (r["nfac"] >= 2) & np.isfinite(r["F"]) & np.isfinite(r["F2"]) & (r["F"] > 0)
        ratio = r["F2"][ok] / (r["df1"][ok] * r["F"][ok])
        assert ratio.max() <= 0.25 * (1 + 1e-9)
        worst = max(worst, float(ratio.max()))
    assert worst > 0.2  # the bound is nearly attained, so it is not loose by construction


def test_oracle_counts_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(5, 60))
        ga = rng.integers(0, 4, n).astype(np.int8)
        gb = rng.integers(0, 4, n).astype(np.int8)
        y = rng.normal(size=n).astype(np.float32)
        y[rng.random(n) < 0.1] = np.nan
        y[rng.random(n) < 0.05] = -9
        r = one_pair(ga, gb, y)
        cnt = np.zeros(9, int)
        sm = np.zeros(9)
        for k in range(n):
            if ga[k] == 3 or gb[k] == 3 or not np.isfinite(y[k]) or y[k] == -9:
                continue
            cnt[3 * ga[k] + gb[k]] += 1
            sm[3 * ga[k] + gb[k]] += float(y[k])
        assert (r["counts"][0] == cnt).all()
        assert np.allclose(r["sums"][0], sm, rtol=0, atol=1e-12)


def test_oracle_known_answer():
    """Hand-computed: cells {0: [1, 3], 4: [10, 12], 8: [5]}; 8df F by hand."""
    ga = [0, 0, 1, 1, 2]
    gb = [0, 0, 1, 1, 2]
    y = [1, 3, 10, 12, 5]
    r = one_pair(ga, gb, y)
    # means 2, 11, 5; grand 6.2; SSB = 2*4.2^2 + 2*4.8^2 + 1.2^2 = 82.8; SSW = 2 + 2 = 4
    # d1 = 2, d2 = 2 -> F = (82.8/2)/(4/2) = 20.7
    assert r["nfac"][0] == 3 and r["df1"][0] == 2 and r["df2"][0] == 2
    assert abs(r["F"][0] - 20.7) < 1e-12


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_invariances(seed):
    rng = np.random.default_rng(seed)
    n = 400
    ga = rng.binomial(2, 0.3, n).astype(np.int8)
    gb = rng.binomial(2, 0.45, n).astype(np.int8)
    ga[rng.random(n) < 0.03] = 3
    # on a 2^-10 grid so that the affine map below is exact in FP32
    y = (np.round((rng.normal(size=n) + 0.3 * (ga == 2)) * 1024) / 1024).astype(np.float32)
    y[rng.random(n) < 0.02] = np.nan
    base = one_pair(ga, gb, y)
    F, F2 = base["F"][0], base["F2"][0]
    # 1. swap roles (transpose the table)
    r = one_pair(gb, ga, y)
    assert rel_err(r["F"][0], F) < 1e-12 and rel_err(r["F2"][0], F2) < 1e-12
    # 2. relabel 0 <-> 2 on one SNP
    ga2 = ga.copy()
    ga2[ga == 0], ga2[ga == 2] = 2, 0
    r = one_pair(ga2, gb, y)
    assert rel_err(r["F"][0], F) < 1e-12 and rel_err(r["F2"][0], F2) < 1e-12
    # 3. affine phenotype (exact in FP32 for these factors)
    r = one_pair(ga, gb, (y * np.float32(4.0) - np.float32(3.0)).astype(np.float32))
    assert rel_err(r["F"][0], F) < 1e-11 and rel_err(r["F2"][0], F2) < 1e-11
    # 4. permute individuals jointly
    p = rng.permutation(n)
    r = one_pair(ga[p], gb[p], y[p])
    assert rel_err(r["F"][0], F) < 1e-12 and rel_err(r["F2"][0], F2) < 1e-12
    # 5. adding excluded individuals changes nothing
    r = one_pair(np.concatenate([ga, [3, 1, 0]]), np.concatenate([gb, [0, 3, 1]]),
                 np.concatenate([y, [5.0, 7.0, -9.0]]).astype(np.float32))
    assert rel_err(r["F"][0], F) < 1e-12 and rel_err(r["F2"][0], F2) < 1e-12
    # sum of counts = complete cases; all nine cells -> df1 = 8, d2 = N' - 9
    ok = (ga != 3) & (gb != 3) & np.isfinite(y)
    assert base["counts"][0].sum() == ok.sum()
    if base["nfac"][0] == 9:
        assert base["df1"][0] == 8 and base["df2"][0] == ok.sum() - 9


def test_oracle_scan_pair_set_and_blocks():
    """Every unordered pair once, in reference block order."""
    nid, sizes = 50, [7, 5, 6]
    words, y = make_dataset(4, nid, sum(sizes))
    codes = O.unpack(words, nid)
    n, hits = O.scan(codes, y, [0, 7, 12], sizes, "f", t1=-1.0)  # every defined pair is a hit
    m = sum(sizes)
    assert n == m * (m - 1) // 2
    assert (hits["snp1"] > hits["snp2"]).all()
    chr_of = np.repeat(np.arange(3), sizes)
    i, j = chr_of[hits["snp1"]], chr_of[hits["snp2"]]
    assert (hits["block"] == i * (i + 1) // 2 + j).all()
    n2, _ = O.scan(codes, y, [0, 7, 12], sizes, "f", t1=-1.0, blk_begin=3, blk_end=4)  # (2, 0)
    assert n2 == 6 * 7
    n3, _ = O.scan(codes, y, [0, 7, 12], sizes, "f", t1=-1.0, blk_begin=5, blk_end=6)  # (2, 2)
    assert n3 == 6 * 5 // 2


def test_oracle_hit_rules():
    """'i' requires nfac > 6; strict '>' on thresholds."""
    nid = 300
    words, y = make_dataset(8, nid, 40, maf=(0.02, 0.08))
    codes = O.unpack(words, nid)
    aa, bb = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    sel = aa > bb
    pr = np.stack([aa[sel], bb[sel]], 1)
    r = O.pairs(codes, y, pr)
    n_f, hf = O.scan(codes, y, [0], [40], "f", t1=0.5)
    n_i, hi = O.scan(codes, y, [0], [40], "i", t1=0.5, t2=0.5)
    ok = r["df1"] >= 1
    assert n_f == int(((r["F"] > np.float32(0.5)) & ok).sum())
    exp_i = ((r["F"] > 0.5) | (r["F2"] > 0.5)) & (r["nfac"] > 6) & ok
    assert n_i == int(exp_i.sum())
    assert (r["nfac"] <= 6).any()


def test_oracle_safe_mode_nine_is_a_value():
    """Safe mode (-s 1): -9 is an ordinary phenotype value, so F is the one-way ANOVA over the cells with the
    -9 individuals included; in normal mode they are excluded. Pinned by scipy's f_oneway on the groups."""
    from scipy.stats import f_oneway
    rng = np.random.default_rng(77)
    for trial in range(20):
        nid = int(rng.integers(40, 200))
        ga = rng.integers(0, 3, nid).astype(np.int8)
        gb = rng.integers(0, 3, nid).astype(np.int8)
        ga[rng.choice(nid, 3, replace=False)] = 3
        y = rng.normal(size=nid).astype(np.float32)
        nine = rng.choice(nid, 6, replace=False)
        y[nine] = -9.0
        y[rng.choice(nid, 2, replace=False)] = np.nan
        codes = np.stack([ga, gb])
        pr = np.array([[0, 1]], np.int32)
        for safe in (False, True):
            r = O.pairs(codes, y, pr, safe=safe)
            keep = (ga != 3) & (gb != 3) & ~np.isnan(y) & (safe | (y != -9.0))
            cells = 3 * ga[keep] + gb[keep]
            groups = [y[keep][cells == c].astype(np.float64) for c in range(9) if (cells == c).any()]
            assert r["counts"][0].sum() == keep.sum()
            if len(groups) >= 2 and keep.sum() > len(groups):
                F = f_oneway(*groups).statistic
                assert abs(r["F"][0] - F) <= 1e-9 * max(1.0, abs(F))


def test_oracle_safe_mode_hit_rule():
    """Safe test 'i' has no nfac > 6 condition (src/EpiOpenCL.c:949-950); test 'f' is unchanged."""
    nid = 300
    words, y = make_dataset(8, nid, 40, maf=(0.02, 0.08))
    codes = O.unpack(words, nid)
    aa, bb = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    sel = aa > bb
    pr = np.stack([aa[sel], bb[sel]], 1)
    r = O.pairs(codes, y, pr, safe=True)
    ok = r["df1"] >= 1
    n_i, _ = O.scan(codes, y, [0], [40], "i", t1=0.5, t2=0.5, safe=True)
    assert n_i == int((((r["F"] > 0.5) | (r["F2"] > 0.5)) & ok).sum())
    n_f, _ = O.scan(codes, y, [0], [40], "f", t1=0.5, safe=True)
    assert n_f == int(((r["F"] > np.float32(0.5)) & ok).sum())
    n_i0, _ = O.scan(codes, y, [0], [40], "i", t1=0.5, t2=0.5)
    assert n_i > n_i0  # pairs with nfac <= 6 now count


def test_ref32_close_to_oracle():
    """The FP32 reference-shaped baseline tracks the FP64 oracle (its FP32 error is not a parity target)."""
    nid, nsnp = 1000, 60
    words, y = make_dataset(21, nid, nsnp, gmiss=0.01, pmiss=0.02)
    codes = O.unpack(words, nid)
    aa, bb = np.meshgrid(np.arange(nsnp), np.arange(nsnp), indexing="ij")
    sel = aa > bb
    pr = np.stack([aa[sel], bb[sel]], 1)
    r = O.pairs(codes, y, pr)
    r32 = O.ref32_pairs(words, nid, y, pr, "i")
    ok = r["df1"] >= 1
    assert (r32["df1"][ok] == r["df1"][ok]).all() and (r32["df2"][ok] == r["df2"][ok]).all()
    assert rel_err(r32["F"][ok], r["F"][ok]).max() < 1e-3


# ----------------------------------------------------------- codec pins --

def test_codec_known_answer():
    import epigpu_amd as E
    codes = np.array([[0, 1, 2, 3] * 4 + [2, 1, 3]], np.int8)  # nid = 19, remain 3
    w = E.pack(codes)
    assert w.shape == (1, 2)
    assert np.uint32(w[0, 0]) == 0x1B1B1B1B
    assert w[0, 1] == 0b100111
    assert (E.unpack(w, 19) == codes).all()
    assert (O.unpack(w, 19) == codes).all()


@pytest.mark.parametrize("nid", [16, 17, 20, 31, 500, 4001])
def test_codec_roundtrip(nid):
    import epigpu_amd as E
    rng = np.random.default_rng(nid)
    codes = rng.integers(0, 4, (13, nid)).astype(np.int8)
    w = E.pack(codes)
    assert w.shape == (13, E.layout(nid)[2])
    assert (E.unpack(w, nid) == codes).all()
    assert (O.unpack(w, nid) == codes).all()


def _ref():
    L = O.ref_readplink()
    if L is None:
        pytest.skip("reference readplink.c not built (oracle/_ref)")
    return L


@pytest.mark.parametrize("nid", [16, 17, 500, 4003])
def test_codec_matches_reference_pack(nid):
    """The reference's own pack()/unpack() (src/readplink.c:1066-1143), compiled from its source."""
    import epigpu_amd as E
    L = _ref()
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(nid + 1)
    nsnp = 9
    codes = rng.integers(0, 4, (nsnp, nid)).astype(np.int8)
    npack, remain = ctypes.c_int(), ctypes.c_int()
    gp = ctypes.POINTER(ctypes.c_int)()
    L.pack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                       ctypes.POINTER(ctypes.POINTER(ctypes.c_int)), ctypes.c_void_p]
    L.pack(nid, nsnp, ctypes.byref(npack), ctypes.byref(remain), ctypes.byref(gp), codes.ctypes.data)
    npack1 = npack.value + (remain.value > 0)
    ref_words = np.ctypeslib.as_array(gp, shape=(nsnp * npack1,)).copy().reshape(nsnp, npack1)
    assert (npack.value, remain.value) == (nid // 16, nid % 16)
    assert (E.pack(codes) == ref_words).all()
    gout = ctypes.c_void_p()
    L.unpack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                         ctypes.POINTER(ctypes.c_void_p)]
    L.unpack(nid, nsnp, npack.value, remain.value, ctypes.cast(gp, ctypes.c_void_p), ctypes.byref(gout))
    ref_codes = np.ctypeslib.as_array(ctypes.cast(gout, ctypes.POINTER(ctypes.c_int8)), shape=(nsnp * nid,))
    assert (ref_codes.reshape(nsnp, nid) == codes).all()
    assert (O.unpack(ref_words, nid) == codes).all()
    libc.free(ctypes.cast(gp, ctypes.c_void_p))
    libc.free(gout)
