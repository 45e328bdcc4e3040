"""PLINK ingest (the reference's data mode, src/readplink.c:1296-1435): text .ped/.map,
PLINK 1 binary .bed/.bim/.fam, cleaning, imputation, random phenotype and extraction
back to text. CPU only.

Pins: a hand-derived known answer; the committed fixtures under tests/golden/ that the
reference's own readplink.c wrote (make_plink_golden.py); and, where the reference
library is built here (oracle/_ref), byte-for-byte comparison of whole containers and
extracted text on further seeded datasets."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import epigpu_amd as E
import oracle as O
from epigpu_amd import scan as S
from plink_data import make_dataset, write_bed, write_ped_map

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(S.LIBDIR, "episcan")

SMALL_PED = """F1 I1 0 0 1 0.5 A A C G 0 0
F2 I2 0 0 2 -1.25 A G C C T T
F3 I3 0 0 1 2.0 G G G G 0 T
F4 I4 0 0 2 -9 G A C C A T
"""
SMALL_MAP = "1 rs1 0 100\n1 rs2 0.5 200\n2 rs3 0 300\n"


def _need_ref():
    if O.ref_readplink() is None:
        pytest.skip("reference readplink.c not built (oracle/_ref)")


def _bytes(p):
    return open(p, "rb").read()


def test_known_answer(tmp_path):
    """Classes and allele names worked out by hand from readped's rules: the first homozygote
    in individual order is class 0; allele names are the last heterozygote's pair, put in
    class order; a call whose first allele is '0' is missing."""
    (tmp_path / "a.ped").write_text(SMALL_PED)
    (tmp_path / "a.map").write_text(SMALL_MAP)
    d = E.read_plink((str(tmp_path / "a.ped"), str(tmp_path / "a.map")))
    assert (d["nid"], d["nsnp"], d["nchr"]) == (4, 3, 2)
    # SNP1 AA AG GG GA: first hom A -> 0; last het "G A" reordered to "A G"
    # SNP2 CG CC GG CC: first hom C -> 0; het "C G" already in class order
    # SNP3 00 TT 0T AT: "0 T" is missing; first hom T -> 0; het "A T" -> "T A"
    assert d["codes"].tolist() == [[0, 1, 2, 1], [1, 0, 2, 0], [3, 0, 3, 1]]
    assert d["map"]["allelename"].tolist() == [b"AG", b"CG", b"TA"]
    assert d["chr"][["chrid", "chrname", "chrsize", "chrstart"]].tolist() == [(0, b"1", 2, 0), (1, b"2", 1, 2)]
    assert d["ped"]["phen"].tolist() == [0.5, -1.25, 2.0, -9.0]
    assert d["ped"]["sex"].tolist() == [b"\x01", b"\x02", b"\x01", b"\x02"]
    assert d["map"]["pd"].tolist() == [100, 200, 300] and d["map"]["gd"].tolist() == [0.0, 0.5, 0.0]
    # packed words follow layout A1 (checked against the codec)
    assert (d["words"] == E.pack(d["codes"].astype(np.int8))).all()
    # SNP without heterozygotes: names are {class-0 allele, 0}
    (tmp_path / "b.ped").write_text("a b 0 0 1 1.5 C C\nc d 0 0 1 2.5 G G\n")
    (tmp_path / "b.map").write_text("1 s 0 1\n")
    d = E.read_plink((str(tmp_path / "b.ped"), str(tmp_path / "b.map")))
    assert d["codes"].tolist() == [[0, 2]] and d["map"]["allelename"].tolist() == [b"C"]


def test_clean_known_answer(tmp_path):
    """cleangeno's thresholds (src/readplink.c:417-497): allele-2 frequency > 0.01, call rate
    > 0.9 and position > 0 per SNP; then call rate > 0.9 and phenotype != -9 per individual."""
    nid = 20
    rows = []
    for i in range(nid):
        g = ["A A",                                  # s1 monomorphic: frequency 0 -> dropped
             "A C" if i < 3 else "A A",              # s2 frequency 3/40 -> kept
             "0 0" if i < 2 else ("G G" if i % 2 else "G T"),  # s3 call rate 18/20 = 0.9 -> dropped
             "A C" if i < 3 else "A A",              # s4 position 0 -> dropped
             "0 0" if i == 12 else ("C C" if i else "G G"),    # s5 kept (class-2 frequency 36/38)
             "0 0" if i == 12 else ("T T" if i % 2 else "A T")]  # s6 kept
        ph = "-9" if i == 7 else f"{i * 0.5}"  # i7: phenotype -9 -> dropped
        rows.append(f"f{i} i{i} 0 0 1 {ph} " + " ".join(g))  # i12: 1 of 3 kept SNPs called -> dropped
    (tmp_path / "c.ped").write_text("\n".join(rows) + "\n")
    (tmp_path / "c.map").write_text("1 s1 0 10\n1 s2 0 20\n1 s3 0 30\n2 s4 0 0\n2 s5 0 50\n2 s6 0 60\n")
    d = E.read_plink((str(tmp_path / "c.ped"), str(tmp_path / "c.map")), clean=True)
    assert d["map"]["snpname"].tolist() == [b"s2", b"s5", b"s6"]
    assert d["report"] == {"snps_in": 6, "snps_kept": 3, "ids_in": 20, "ids_kept": 18, "imputed": 0}
    ids = d["ped"]["id"].tolist()
    assert b"i7" not in ids and b"i12" not in ids and len(ids) == 18
    assert d["codes"][1].tolist() == [0] + [2] * 17  # s5: G G first -> class 0
    # both chromosomes stay listed, sizes and starts over the kept SNPs
    assert d["chr"][["chrname", "chrsize", "chrstart"]].tolist() == [(b"1", 1, 0), (b"2", 2, 1)]


def test_fixture_containers_match_reference_output():
    """The reference's -Dr and -Drcmn containers for the committed fixture, byte for byte."""
    ped, mp = os.path.join(GOLD, "plink_small.ped"), os.path.join(GOLD, "plink_small.map")
    for name, kw in (("plink_small.epigpu", {}), ("plink_small_cmn.epigpu", dict(clean=True, impute=True,
                                                                                  rnorm=True))):
        exp = E.read_container(os.path.join(GOLD, name))
        got = E.read_plink((ped, mp), **kw)
        assert got["nid"] == exp["nid"] and got["nsnp"] == exp["nsnp"] and got["nchr"] == exp["nchr"]
        assert (got["words"] == exp["geno"]).all(), name
        assert got["ped"].tobytes() == exp["ped"].tobytes(), name
        assert got["map"].tobytes() == exp["map"].tobytes(), name
        assert got["chr"].tobytes() == exp["chr"].tobytes(), name


def test_fixture_container_file_and_extract(tmp_path):
    """Written container file and extracted text equal the reference's files."""
    ped, mp = os.path.join(GOLD, "plink_small.ped"), os.path.join(GOLD, "plink_small.map")
    out = str(tmp_path / "o.epigpu")
    E.read_plink((ped, mp), clean=True, impute=True, rnorm=True, container=out)
    assert _bytes(out) == _bytes(os.path.join(GOLD, "plink_small_cmn.epigpu"))
    E.write_ped_map(out, str(tmp_path / "x.ped"), str(tmp_path / "x.map"))
    assert _bytes(tmp_path / "x.ped") == _bytes(os.path.join(GOLD, "plink_small_cmn.x.ped"))
    assert _bytes(tmp_path / "x.map") == _bytes(os.path.join(GOLD, "plink_small_cmn.x.map"))


@pytest.mark.parametrize("seed", [3, 19])
@pytest.mark.parametrize("flags", ["", "c", "m", "n", "cm", "cmn"])
def test_containers_match_reference_data_mode(tmp_path, capfd, seed, flags):
    _need_ref()
    ds = make_dataset(seed)
    ped, mp = str(tmp_path / "d.ped"), str(tmp_path / "d.map")
    write_ped_map(ds, ped, mp, zero_first=0.3)
    kw = dict(clean="c" in flags, impute="m" in flags, rnorm="n" in flags)
    O.ref_data_mode(ped, mp, str(tmp_path / "ref.epigpu"), **kw)
    d = E.read_plink((ped, mp), container=str(tmp_path / "our.epigpu"), **kw)
    capfd.readouterr()
    assert _bytes(tmp_path / "our.epigpu") == _bytes(tmp_path / "ref.epigpu")
    if "m" in flags:
        assert d["report"]["imputed"] > 0 and (d["codes"] != 3).all()


def test_extract_matches_reference(tmp_path, capfd):
    """-De text, for a continuous and a 1/2-coded phenotype."""
    _need_ref()
    ds = make_dataset(5)
    write_ped_map(ds, str(tmp_path / "d.ped"), str(tmp_path / "d.map"))
    c = E.read_plink((str(tmp_path / "d.ped"), str(tmp_path / "d.map")), clean=True)
    for binary in (False, True):
        ped = c["ped"].copy()
        if binary:
            ped["phen"] = np.where(np.arange(c["nid"]) % 3 == 0, 2.0, 1.0)
        src = str(tmp_path / f"s{binary}.epigpu")
        E.write_container(src, c["words"], c["nid"], ped, c["map"], c["chr"])
        O.ref_extract(src, str(tmp_path / "r.ped"), str(tmp_path / "r.map"))
        E.write_ped_map(src, str(tmp_path / "o.ped"), str(tmp_path / "o.map"))
        capfd.readouterr()
        assert _bytes(tmp_path / "o.ped") == _bytes(tmp_path / "r.ped")
        assert _bytes(tmp_path / "o.map") == _bytes(tmp_path / "r.map")


@pytest.mark.parametrize("flags", ["", "cm"])
def test_bed_equals_ped(tmp_path, capfd, flags):
    """A .bed fileset reads as the .ped PLINK writes for the same data (heterozygotes "A1 A2",
    missing "0 0"), and, where built, as the reference reads that .ped."""
    ds = make_dataset(41)
    ds["g"][:, 9] = np.where(ds["g"][:, 9] < 0, -1, 2)  # no A1 copies: A1 may be PLINK's "0"
    write_ped_map(ds, str(tmp_path / "d.ped"), str(tmp_path / "d.map"), het_order="a1a2")
    write_bed(ds, str(tmp_path / "d"), a1_override={9: "0"})
    kw = dict(clean="c" in flags, impute="m" in flags)
    a = E.read_plink(str(tmp_path / "d"), container=str(tmp_path / "bed.epigpu"), **kw)
    b = E.read_plink((str(tmp_path / "d.ped"), str(tmp_path / "d.map")), container=str(tmp_path / "ped.epigpu"),
                     **kw)
    assert (a["codes"] == b["codes"]).all()
    assert _bytes(tmp_path / "bed.epigpu") == _bytes(tmp_path / "ped.epigpu")
    if O.ref_readplink() is not None:
        O.ref_data_mode(str(tmp_path / "d.ped"), str(tmp_path / "d.map"), str(tmp_path / "ref.epigpu"), **kw)
        capfd.readouterr()
        assert _bytes(tmp_path / "bed.epigpu") == _bytes(tmp_path / "ref.epigpu")


def test_tab_separated_and_round_trip(tmp_path):
    """Tab-separated text reads like space-separated text, and extract -> read gives back the
    same classes, records and container."""
    ds = make_dataset(8, awkward=False)
    write_ped_map(ds, str(tmp_path / "s.ped"), str(tmp_path / "s.map"))
    write_ped_map(ds, str(tmp_path / "t.ped"), str(tmp_path / "t.map"), sep="\t")
    a = E.read_plink((str(tmp_path / "s.ped"), str(tmp_path / "s.map")), container=str(tmp_path / "s.epigpu"))
    b = E.read_plink((str(tmp_path / "t.ped"), str(tmp_path / "t.map")), container=str(tmp_path / "t.epigpu"))
    assert _bytes(tmp_path / "s.epigpu") == _bytes(tmp_path / "t.epigpu")
    assert (a["map"]["allelename"] != b"").all()
    E.write_ped_map(str(tmp_path / "s.epigpu"), str(tmp_path / "x.ped"), str(tmp_path / "x.map"))
    c = E.read_plink((str(tmp_path / "x.ped"), str(tmp_path / "x.map")), container=str(tmp_path / "x.epigpu"))
    assert (c["codes"] == a["codes"]).all()
    assert _bytes(tmp_path / "x.epigpu") == _bytes(tmp_path / "s.epigpu")


@pytest.mark.parametrize("nid", [65, 131])
def test_cleaned_round_trip_relabels_only_leading_class2_snps(tmp_path, nid):
    """After cleaning drops individuals, a SNP's first kept homozygote can be class 2.
    Extracting writes class 0 as the A1 homozygote, so re-reading makes that homozygote
    class 0 again (first homozygote seen, src/readplink.c:266-334): exactly those SNPs
    swap classes 0 and 2; every other SNP, and every missing or het call, is unchanged.
    Sizes cross the 64-individual block boundary of the text reader, and the .bed reader
    gives the same container."""
    ds = make_dataset(7, nid=nid, nsnp=50, chr_sizes=(25, 25))
    write_ped_map(ds, str(tmp_path / "a.ped"), str(tmp_path / "a.map"))
    write_bed(ds, str(tmp_path / "b"))
    a = E.read_plink((str(tmp_path / "a.ped"), str(tmp_path / "a.map")), clean=True,
                     container=str(tmp_path / "a.epigpu"))
    assert (E.read_plink(str(tmp_path / "b"), clean=True)["words"] == a["words"]).all()
    E.write_ped_map(str(tmp_path / "a.epigpu"), str(tmp_path / "x.ped"), str(tmp_path / "x.map"))
    r = E.read_plink((str(tmp_path / "x.ped"), str(tmp_path / "x.map")))
    A, R = a["codes"].astype(int), r["codes"].astype(int)
    lead = np.array([s[(s == 0) | (s == 2)][0] for s in A])
    swapped = np.where((A == 0) | (A == 2), 2 - A, A)
    assert (lead == 2).any()
    assert (R[lead == 2] == swapped[lead == 2]).all()
    assert (R[lead == 0] == A[lead == 0]).all()


def test_random_steps_follow_srand_stream(tmp_path):
    """Imputation and the random phenotype draw the sequence rand() gives after srand(seed),
    without disturbing the process's own rand() state."""
    ds = make_dataset(12, awkward=False)
    write_ped_map(ds, str(tmp_path / "d.ped"), str(tmp_path / "d.map"))
    files = (str(tmp_path / "d.ped"), str(tmp_path / "d.map"))
    libc = ctypes.CDLL(None)
    libc.srand(77)
    before = libc.rand()
    libc.srand(77)
    d = E.read_plink(files, rnorm=True, seed=9)
    assert libc.rand() == before
    libc.srand(9)
    u = []
    for _ in range(2 * d["nid"]):
        u.append(np.float32(libc.rand()) / np.float32(2147483647))
    x1, x2 = np.array(u[0::2], np.float32), np.array(u[1::2], np.float32)
    z = np.sqrt(np.float32(-2) * np.log(x1)) * np.cos((2 * 3.14159 * x2.astype(np.float64)).astype(np.float32))
    assert np.abs(d["ped"]["phen"] - z).max() <= 2e-6 * max(1.0, np.abs(z).max())
    assert d["ped"]["phen"].tolist() != E.read_plink(files, rnorm=True, seed=10)["ped"]["phen"].tolist()


BAD = {
    "ragged": ("a b 0 0 1 1 A A C C\nc d 0 0 1 1 A A\n", "1 s1 0 1\n1 s2 0 2\n"),
    "third_allele": ("a b 0 0 1 1 A A\nc d 0 0 1 1 C C\ne f 0 0 1 1 G G\n", "1 s1 0 1\n"),
    "split_chr": ("a b 0 0 1 1 A A C C G G\nc d 0 0 1 1 A C C G G T\n", "1 s1 0 1\n2 s2 0 2\n1 s3 0 3\n"),
    "map_short": ("a b 0 0 1 1 A A C C\nc d 0 0 1 1 A C C G\n", "1 s1 0 1\n"),
    "map_cols": ("a b 0 0 1 1 A A\nc d 0 0 1 1 A C\n", "1 s1 1\n"),
    "long_name": ("a" * 30 + " b 0 0 1 1 A A\nc d 0 0 1 1 A C\n", "1 s1 0 1\n"),
    "phen_na": ("a b 0 0 1 NA A A\nc d 0 0 1 1 A C\n", "1 s1 0 1\n"),
    "two_char_allele": ("a b 0 0 1 1 AG AG\nc d 0 0 1 1 A C\n", "1 s1 0 1\n"),
    "odd_columns": ("a b 0 0 1 1 A A C\nc d 0 0 1 1 A C C\n", "1 s1 0 1\n"),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_malformed_text_is_refused(tmp_path, case):
    ped, mp = BAD[case]
    (tmp_path / "d.ped").write_text(ped)
    (tmp_path / "d.map").write_text(mp)
    with pytest.raises(S.EpiError) as e:
        E.read_plink((str(tmp_path / "d.ped"), str(tmp_path / "d.map")))
    assert e.value.code == -102
    assert str(e.value)


def test_malformed_bed_is_refused(tmp_path):
    ds = make_dataset(2, nid=10, nsnp=10, chr_sizes=(10,), awkward=False)
    pre = str(tmp_path / "d")
    write_bed(ds, pre)
    raw = _bytes(pre + ".bed")
    for bad in (b"\x6c\x1b\x00" + raw[3:], b"XY\x01" + raw[3:], raw[:-1], raw + b"\x00"):
        open(pre + ".bed", "wb").write(bad)
        with pytest.raises(S.EpiError) as e:
            E.read_plink(pre)
        assert e.value.code == -102
    os.remove(pre + ".fam")
    with pytest.raises(S.EpiError) as e:
        E.read_plink(pre)
    assert e.value.code == -101


def test_impute_refuses_snp_without_calls(tmp_path):
    (tmp_path / "d.ped").write_text("a b 0 0 1 1 A A 0 0\nc d 0 0 1 2 A C 0 0\n")
    (tmp_path / "d.map").write_text("1 s1 0 1\n1 s2 0 2\n")
    with pytest.raises(S.EpiError):
        E.read_plink((str(tmp_path / "d.ped"), str(tmp_path / "d.map")), impute=True)


def test_cli_data_mode(tmp_path):
    ped, mp = os.path.join(GOLD, "plink_small.ped"), os.path.join(GOLD, "plink_small.map")
    out = str(tmp_path / "o.epigpu")
    r = subprocess.run([EXE, "-Drcmn", ped, mp, out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert _bytes(out) == _bytes(os.path.join(GOLD, "plink_small_cmn.epigpu"))
    r = subprocess.run([EXE, "-De", str(tmp_path / "x.ped"), str(tmp_path / "x.map"), out], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    assert _bytes(tmp_path / "x.ped") == _bytes(os.path.join(GOLD, "plink_small_cmn.x.ped"))
    ds = make_dataset(41)
    write_bed(ds, str(tmp_path / "b"))
    r = subprocess.run([EXE, "-Dbc", str(tmp_path / "b"), str(tmp_path / "b.epigpu")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    E.read_plink(str(tmp_path / "b"), clean=True, container=str(tmp_path / "lib.epigpu"))
    assert _bytes(tmp_path / "b.epigpu") == _bytes(tmp_path / "lib.epigpu")
    for bad in (["-Drx", ped, mp, out], ["-Dr", ped, out], ["-Dre", ped, mp, out], ["-Dem", ped, mp, out]):
        assert subprocess.run([EXE, *bad], capture_output=True).returncode == 1
    r = subprocess.run([EXE, "-Dr", str(tmp_path / "absent.ped"), mp, out], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
