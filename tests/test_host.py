"""Host side (CPU): C ABI exports, F-distribution helper, container format, output
format, resume parsing and the synthetic generator. No GPU compute."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import epigpu_amd as E
from epigpu_amd import scan as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", txt, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("episcan.h", "libepiscan.so"), ("epihost.h", "libepihost.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared(header)
    assert len(names) > 10
    L = ctypes.CDLL(os.path.join(S.LIBDIR, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_compat_entry_exported():
    L = ctypes.CDLL(os.path.join(S.LIBDIR, "libepihost.so"))
    for n in ("perform_scan", "UTEST", "UTHRESHOLD1", "UTHRESHOLD2", "UDEVICE", "UPERM", "plinkbin", "UPHEN"):
        assert hasattr(L, n)


def test_open_without_device_fails_cleanly():
    L = S.episcan()
    n = ctypes.c_int()
    ctx = ctypes.c_void_p()
    rc = L.epi_open(10_000, ctypes.byref(ctx))
    assert rc == -3 and not ctx.value  # EPI_ERR_NODEV
    assert L.epi_strerror(rc) == b"no such device"


def test_f_isf_matches_scipy_fixture():
    g = json.load(open(os.path.join(GOLD, "fdist.json")))
    for c in g["isf"]:
        v = E.f_isf(c["p"], c["d1"], c["d2"])
        # the survival function is pinned to scipy below; the inverse must round-trip
        assert abs(E.f_sf(v, c["d1"], c["d2"]) - c["p"]) <= 1e-10 * c["p"], c
        if abs(c["sf_at_isf"] - c["p"]) <= 1e-10 * c["p"]:  # scipy's inverse is tight here
            assert abs(v - c["isf"]) <= 1e-8 * c["isf"], c
    for c in g["sf"]:
        v = E.f_sf(c["F"], c["d1"], c["d2"])
        assert abs(v - c["sf"]) <= 1e-10 * max(c["sf"], 1e-300) + 1e-300, c


def test_f_thresholds_of_the_workloads():
    assert round(E.f_isf(1e-4, 8, 491), 6) == 4.084638
    assert round(E.f_isf(1e-4, 8, 3991), 6) == 3.991350
    assert round(E.f_isf(1e-4, 4, 3991), 6) == 5.894055


def test_betai_closed_forms():
    # I_x(1, 1) = x; I_x(a, 1) = x^a; I_x(1, b) = 1 - (1-x)^b
    for x in (0.01, 0.3, 0.77, 0.999):
        assert abs(S.betai(1, 1, x) - x) < 1e-14
        assert abs(S.betai(3.5, 1, x) - x ** 3.5) < 1e-13
        assert abs(S.betai(1, 2.5, x) - (1 - (1 - x) ** 2.5)) < 1e-13


def test_format_row_reference_strings():
    # src/EpiOpenCL.c:1632-1653: 1-based global indices, %f floats
    assert S.format_row("f", 12, "rs13", "1", 3, "rs4", "1", 8, 491, 7.25) == "13\trs13\t1\t4\trs4\t1\t8\t491\t7.250000\n"
    assert S.format_row("i", 0, "a", "2", 1, "b", "1", 7, 100, 1.5, 11.0) == \
        "1\ta\t2\t2\tb\t1\t7\t100\t1.500000\t11.000000\n"


# ------------------------------------------------------------ container --

def small_dataset(nid=37, sizes=(5, 4)):
    nsnp = sum(sizes)
    words = E.synth_genotypes(9, nid, nsnp, 0.1, 0.5, 0.05)
    y = E.synth_phenotype(9, 0, nid)
    ped, mp, chrs = S.make_records(nid, nsnp, list(sizes), y)
    return words, ped, mp, chrs


def test_container_roundtrip(tmp_path):
    words, ped, mp, chrs = small_dataset()
    p = str(tmp_path / "d.epigpu")
    E.write_container(p, words, 37, ped, mp, chrs)
    assert os.path.getsize(p) == 6 + 20 + 128 * 37 + 88 * 9 + 44 * 2 + 4 * 3 * 9
    c = E.read_container(p)
    assert (c["geno"] == words).all() and c["nid"] == 37 and c["nsnp"] == 9 and c["nchr"] == 2
    assert c["ped"].tobytes() == ped.tobytes() and c["map"].tobytes() == mp.tobytes()
    assert c["chr"].tobytes() == chrs.tobytes()


def test_container_rejects_bad_files(tmp_path):
    words, ped, mp, chrs = small_dataset()
    p = str(tmp_path / "d.epigpu")
    E.write_container(p, words, 37, ped, mp, chrs)
    raw = open(p, "rb").read()
    bad = tmp_path / "bad.epigpu"
    bad.write_bytes(b"notGPU" + raw[6:])
    with pytest.raises(S.EpiError):
        E.read_container(str(bad))
    bad.write_bytes(raw[:-5])  # truncated words
    with pytest.raises(S.EpiError):
        E.read_container(str(bad))
    with pytest.raises(S.EpiError):
        E.read_container(str(tmp_path / "absent"))


def _ref():
    import oracle as O
    L = O.ref_readplink()
    if L is None:
        pytest.skip("reference readplink.c not built (oracle/_ref)")
    return L


def test_container_matches_reference_reader_and_writer(tmp_path, capfd):
    """Byte compatibility with src/readplink.c:1145-1273, compiled from the reference source."""
    L = _ref()
    words, ped, mp, chrs = small_dataset()
    ours = str(tmp_path / "ours.epigpu")
    E.write_container(ours, words, 37, ped, mp, chrs)
    # reference reader on our file
    nid, nsnp, npack, remain, nchr = (ctypes.c_int() for _ in range(5))
    gm, dt, gp, cs = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.readpackedbinary(ctypes.byref(nid), ctypes.byref(nsnp), ctypes.byref(npack), ctypes.byref(remain),
                       ctypes.byref(nchr), ctypes.byref(gm), ctypes.byref(dt), ctypes.byref(gp), ctypes.byref(cs),
                       ours.encode())
    assert (nid.value, nsnp.value, npack.value, remain.value, nchr.value) == (37, 9, 2, 5, 2)
    got = np.ctypeslib.as_array(ctypes.cast(gp, ctypes.POINTER(ctypes.c_int32)), shape=(27,))
    assert (got.reshape(9, 3) == words).all()
    assert ctypes.string_at(dt.value, 128 * 37) == ped.tobytes()
    assert ctypes.string_at(gm.value, 88 * 9) == mp.tobytes()
    # reference writer -> our reader
    theirs = str(tmp_path / "theirs.epigpu")
    w = np.ascontiguousarray(words)
    L.createpackedbinary(37, 9, 2, 5, 2, ctypes.c_void_p(mp.ctypes.data), ctypes.c_void_p(ped.ctypes.data),
                         ctypes.c_void_p(w.ctypes.data), ctypes.c_void_p(chrs.ctypes.data), theirs.encode())
    assert open(theirs, "rb").read() == open(ours, "rb").read()
    capfd.readouterr()


# --------------------------------------------------------------- resume --

HDR = ("Datafile: data.epigpu \nNumber of individuals: 100\nNumber of SNPs: 50\nPermutation: 0\n"
       "Threshold: 6.500000 10.500000\nTest: interactions + full\n\n"
       "SNP1\tSNP1name\tChr1\tSNP2\tSNP2name\tChr2\tdf1\tdf2\tFval\tFint\n")
ROW = "2\trs2\t1\t1\trs1\t1\t8\t91\t7.000000\t1.000000\n"


def test_resume_states(tmp_path):
    f = tmp_path / "out.txt"
    st = S.resume_state(str(f), "data.epigpu", 100, 50, test="i")
    assert st["exists"] == 0 and (st["chr1"], st["chr2"]) == (0, 0)
    f.write_text(HDR)
    st = S.resume_state(str(f), "data.epigpu", 100, 50, test="i")
    assert st["exists"] == 1 and st["blocks"] == 0 and (st["chr1"], st["chr2"]) == (0, 0)
    f.write_text(HDR + "# 0 x 0 : 2 interactions\n" + ROW * 2 + "# 1 x 0 : 0 interactions\n")
    st = S.resume_state(str(f), "data.epigpu", 100, 50, test="i")
    assert (st["chr1"], st["chr2"], st["tothits"], st["blocks"]) == (1, 1, 2, 2)
    f.write_text(HDR + "# 0 x 0 : 1 interactions\n" + ROW + "# 1 x 0 : 0 interactions\n# 1 x 1 : 1 interactions\n"
                 + ROW)
    st = S.resume_state(str(f), "data.epigpu", 100, 50, test="i")
    assert (st["chr1"], st["chr2"], st["tothits"]) == (2, 0, 2)


def test_resume_drops_cut_block(tmp_path):
    f = tmp_path / "out.txt"
    good = HDR + "# 0 x 0 : 1 interactions\n" + ROW
    f.write_text(good + "# 1 x 0 : 3 interactions\n" + ROW + ROW[:7])
    st = S.resume_state(str(f), "data.epigpu", 100, 50, test="i")
    assert (st["chr1"], st["chr2"], st["tothits"], st["blocks"]) == (1, 0, 1, 1)
    assert f.read_text() == good


@pytest.mark.parametrize("kw", [dict(datafile="other.epigpu"), dict(nid=101), dict(nsnp=49), dict(perm=3),
                                dict(test="f")])
def test_resume_refuses_other_analysis(tmp_path, kw):
    f = tmp_path / "out.txt"
    f.write_text(HDR + "# 0 x 0 : 0 interactions\n")
    args = dict(datafile="data.epigpu", nid=100, nsnp=50, perm=0, test="i")
    args.update(kw)
    with pytest.raises(S.EpiError):
        S.resume_state(str(f), args["datafile"], args["nid"], args["nsnp"], args["perm"], args["test"])


# ------------------------------------------------------------ generator --

def test_generator_deterministic_and_shaped():
    w1 = E.synth_genotypes(123, 1001, 300, 0.05, 0.5, 0.01)
    w2 = E.synth_genotypes(123, 1001, 300, 0.05, 0.5, 0.01)
    assert (w1 == w2).all()
    assert not (E.synth_genotypes(124, 1001, 300, 0.05, 0.5, 0.01) == w1).all()
    c = E.unpack(w1, 1001)
    miss = (c == 3).mean()
    assert 0.008 < miss < 0.012
    obs = np.where(c == 3, np.nan, c.astype(float))
    af = np.nanmean(obs, axis=1) / 2
    assert af.min() > 0.02 and af.max() < 0.56
    y = E.synth_phenotype(5, 0, 200000, S.PHEN_NORMAL)
    assert abs(y.mean()) < 0.01 and abs(y.std() - 1) < 0.01
    g = E.synth_phenotype(5, 1, 200000, S.PHEN_GAMMA2_STD)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.02 and 1.2 < ((g - g.mean()) ** 3).mean() < 1.6
    t = E.synth_phenotype(5, 2, 200000, S.PHEN_T5)
    assert abs(t.var() - 5 / 3) < 0.1
    b = E.synth_phenotype(5, 3, 200000, S.PHEN_BERN02)
    assert set(np.unique(b)) == {0.0, 1.0} and abs(b.mean() - 0.2) < 0.005
    m = E.synth_phenotype(5, 4, 100000, S.PHEN_NORMAL, 0.03)
    assert 0.028 < np.isnan(m).mean() < 0.032


def test_plant_adds_effect_to_one_cell():
    nid, nsnp = 4000, 50
    w = E.synth_genotypes(7, nid, nsnp)
    y0 = E.synth_phenotype(7, 0, nid)
    y = y0.copy()
    pairs = E.synth_plant(7, w, nid, y, 3, 0.15)
    c = E.unpack(w, nid)
    changed = np.zeros(nid, bool)
    for a, b, ga, gb in pairs:
        assert a > b
        changed |= (c[a] == ga) & (c[b] == gb)
    d = y - y0
    assert np.allclose(d[~changed], 0)
    assert (d[changed] > 0.149).all()


def test_permute_phen_is_the_reference_shuffle():
    """Fisher-Yates with the C library's rand() after srand(k), from i = nid down to 2 (src/main.c:336-347),
    restated here with libc's own srand/rand through ctypes."""
    import ctypes.util
    from epigpu_amd import scan as S
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    nid = 257
    ped, _, _ = S.make_records(nid, 4, [4], np.arange(nid, dtype=np.float32) * 0.5 - 3.0)
    for k in (1, 2, 17):
        got = S.permute_phen(k, ped)
        ref = (np.arange(nid, dtype=np.float32) * 0.5 - 3.0).tolist()
        libc.srand(k)
        for i in range(nid, 1, -1):
            r = libc.rand() % i
            ref[r], ref[i - 1] = ref[i - 1], ref[r]
        assert got.tolist() == ref
        assert sorted(got.tolist()) == sorted(ped["phen"].tolist())
    assert not np.array_equal(S.permute_phen(1, ped), S.permute_phen(2, ped))


def test_perform_scan_perms_refuses_existing_outputs(tmp_path):
    """Batched permutations never overwrite or resume: an existing output file is refused before any device work."""
    from epigpu_amd import scan as S
    nid, sizes = 64, [20]
    words = S.synth_genotypes(3, nid, 20)
    ped, mp, chrs = S.make_records(nid, 20, sizes, S.synth_phenotype(3, 0, nid))
    fmt = str(tmp_path / "out.perm%d")
    open(fmt % 2, "w").write("x")
    with pytest.raises(S.EpiError) as e:
        S.perform_scan_perms(words, nid, ped, mp, chrs, fmt, 1, 3)
    assert e.value.code == -105  # EPIH_ERR_DONE
    with pytest.raises(S.EpiError):
        S.perform_scan_perms(words, nid, ped, mp, chrs, str(tmp_path / "bad%s%d"), 1, 2)
    with pytest.raises(S.EpiError):
        S.perform_scan_perms(words, nid, ped, mp, chrs, fmt, 3, 1)


def test_container_rejects_unterminated_names(tmp_path):
    """Name fields are printed with %s, so a container whose name lacks a NUL within its 30 bytes, or whose
    chromosome record points past the SNPs, is refused as malformed rather than read past its end."""
    words, ped, mp, chrs = small_dataset()
    cases = []
    p1 = ped.copy()
    p1[36]["id"] = b"X" * 30
    cases.append((p1, mp, chrs))
    m1 = mp.copy()
    m1[8]["snpname"] = b"r" * 30
    cases.append((ped, m1, chrs))
    c1 = chrs.copy()
    c1[1]["chrname"] = b"c" * 30
    cases.append((ped, mp, c1))
    c2 = chrs.copy()
    c2[1]["chrsize"] = 100
    cases.append((ped, mp, c2))
    for k, (pp, mm, cc) in enumerate(cases):
        path = str(tmp_path / f"bad{k}.epigpu")
        E.write_container(path, words, 37, pp, mm, cc)
        with pytest.raises(S.EpiError) as e:
            E.read_container(path)
        assert e.value.code == -102, k  # EPIH_ERR_FORMAT
    good = str(tmp_path / "good.epigpu")
    p2 = ped.copy()
    p2[0]["id"] = b"Y" * 29  # 29 bytes and the terminator: the longest name
    E.write_container(good, words, 37, p2, mp, chrs)
    assert E.read_container(good)["ped"][0]["id"] == b"Y" * 29


def test_permutation_side_file_is_the_references(tmp_path):
    """perm<k>.txt as permute() writes it (src/main.c:349-356): a "Family ID Phenotype" line, then family, id
    and %f phenotype of every record in its permuted position; the order is the shuffle restated with libc's
    own srand/rand. Phenotypes of a -P file (by original index) replace the records' own."""
    import ctypes.util
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    nid = 41
    y = (np.arange(nid, dtype=np.float32) * 0.25 - 2.0).astype(np.float32)
    ped, _, _ = S.make_records(nid, 4, [4], y)
    for k in (1, 5):
        order = S.permute_order(k, nid)
        ref = list(range(nid))
        libc.srand(k)
        for i in range(nid, 1, -1):
            r = libc.rand() % i
            ref[r], ref[i - 1] = ref[i - 1], ref[r]
        assert order.tolist() == ref
        assert S.permute_phen(k, ped).tolist() == y[ref].tolist()
        path = str(tmp_path / f"perm{k}.txt")
        S.write_perm_file(path, ped, order)
        lines = ["Family ID Phenotype\n"] + ["F%d I%d %f\n" % (j, j, y[j]) for j in ref]
        assert open(path).read() == "".join(lines)
        pfile = (np.arange(nid, dtype=np.float32) * -1.5).astype(np.float32)
        S.write_perm_file(path, ped, order, pfile)
        lines = ["Family ID Phenotype\n"] + ["F%d I%d %f\n" % (j, j, pfile[j]) for j in ref]
        assert open(path).read() == "".join(lines)


def test_perform_scan_multi_refuses_bad_outputs(tmp_path):
    """The batched multi-phenotype driver refuses existing output files and repeated names before any device
    work, so it runs without a GPU."""
    nid, sizes = 64, [20]
    words = S.synth_genotypes(3, nid, 20)
    _, mp, chrs = S.make_records(nid, 20, sizes, S.synth_phenotype(3, 0, nid))
    phen = np.stack([S.synth_phenotype(3, r, nid) for r in range(3)])
    outs = [str(tmp_path / f"o{r}.txt") for r in range(3)]
    open(outs[1], "w").write("x")
    with pytest.raises(S.EpiError) as e:
        S.perform_scan_multi(words, nid, mp, chrs, phen, outs)
    assert e.value.code == -105  # EPIH_ERR_DONE
    with pytest.raises(S.EpiError) as e:
        S.perform_scan_multi(words, nid, mp, chrs, phen, [outs[0], outs[2], outs[0]])
    assert e.value.code == -104  # EPIH_ERR_ARG
