"""GPU tests of phenotype files and permutations through the CLI, the batched multi-phenotype driver and the
reference's own main() linked against libepihost (src/main.c:172-181 -P, :191-222 readphenotype, :309-318 and
:336-357 permute): a -P file's scan against the oracle, several -P files and their permutations in one
batched launch equal to one run each, the perm<k>.txt side files, and the drop-in's thresholds taken as
given (0 included)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import make_dataset
from test_gpu_driver import parse_output

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "epigpu_amd", "lib", "episcan")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "epiGPU-dropin")
SIZES = [150, 110]
NID = 500
T1, T2 = 3.0, 4.0


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    """A container (phenotype N(0,1) with -9 missing) and two phenotype files of other traits."""
    import epigpu_amd as E
    from epigpu_amd import scan as S
    d = tmp_path_factory.mktemp("phen")
    nsnp = sum(SIZES)
    words, y = make_dataset(0x9F1, NID, nsnp, maf=(0.05, 0.5), gmiss=0.01, pmiss=0.02)
    y = np.where(np.isnan(y), np.float32(-9.0), y).astype(np.float32)
    ped, mp, chrs = S.make_records(NID, nsnp, SIZES, y)
    path = str(d / "data.epigpu")
    E.write_container(path, words, NID, ped, mp, chrs)
    files = []
    for j, dist in enumerate((S.PHEN_GAMMA2, S.PHEN_T5)):
        v = E.synth_phenotype(0x9F1, j + 1, NID, dist, 0.0)
        v[np.random.default_rng(j).choice(NID, 9, replace=False)] = -9.0
        f = str(d / f"trait{j + 1}.txt")
        with open(f, "w") as fh:
            fh.write("".join("%.9g\n" % x for x in v))
        files.append((f, v.astype(np.float32)))
    return path, words, ped, mp, chrs, files


def run(args, cwd):
    r = subprocess.run(args, capture_output=True, text=True, timeout=180, cwd=str(cwd))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def hits_of(path):
    _, blocks, rows = parse_output(path)
    return {(a, b): (F, F2) for rs in rows.values() for a, b, d1, d2, F, F2 in rs}, blocks


def test_cli_phenotype_file_matches_oracle(tmp_path, data):
    """-P replaces the container's phenotypes (-9 missing); the hits are the oracle's on the file's values and
    the header names the file."""
    path, words, ped, mp, chrs, files = data
    f, v = files[0]
    out = tmp_path / "out.txt"
    run([EXE, "-A", path, str(out), "-t", "i", "-F", str(T1), "-I", str(T2), "-P", f], tmp_path)
    txt = open(out).read()
    assert txt.startswith("Datafile: data.epigpu %s\n" % f)
    got, blocks = hits_of(out)
    n_ref, ref = O.scan(O.unpack(words, NID), v, [0, SIZES[0]], SIZES, "i", T1, T2)
    want = {(a, b): (F, F2) for a, b, F, F2 in zip(ref["snp1"].tolist(), ref["snp2"].tolist(), ref["F"], ref["F2"])}
    assert n_ref > 50
    for pr in set(got) ^ set(want):  # only threshold ties may differ
        x = ref[(ref["snp1"] == pr[0]) & (ref["snp2"] == pr[1])]
        assert len(x) == 1 and min(abs(x["F"][0] - T1) / T1, abs(x["F2"][0] - T2) / T2) < 1e-5
    for pr in set(got) & set(want):  # %f rows: 6 decimals
        assert abs(got[pr][0] - want[pr][0]) <= 1e-5 * max(1.0, want[pr][0]) + 5e-7
        assert abs(got[pr][1] - want[pr][1]) <= 1e-5 * max(1.0, want[pr][1]) + 5e-7


def test_dropin_phenotype_file_and_permutation(tmp_path, data):
    """The reference's main() with -P and -p 2, linked against libepihost, and this CLI with the same options
    write the same output file and the same perm2.txt."""
    if not os.path.exists(DROPIN):
        pytest.skip("drop-in binary not built (needs the reference tree at build time)")
    path, words, ped, mp, chrs, files = data
    f, v = files[1]
    a, b = tmp_path / "ref", tmp_path / "ours"
    a.mkdir()
    b.mkdir()
    opts = ["-t", "i", "-F", str(T1), "-I", str(T2), "-P", f, "-p", "2"]
    r = run([DROPIN, "-A", path, "out.txt"] + opts, a)
    assert "Permutation 2" in r.stdout
    run([EXE, "-A", path, "out.txt"] + opts, b)
    assert open(a / "out.txt").read() == open(b / "out.txt").read()
    assert open(a / "perm2.txt").read() == open(b / "perm2.txt").read()
    lines = open(b / "perm2.txt").read().splitlines()
    assert lines[0] == "Family ID Phenotype" and len(lines) == NID + 1


def test_several_phenotype_files_one_launch(tmp_path, data):
    """-P f1 -P f2 scans both files in one batched launch into out.P1 and out.P2, each equal to the file of a
    single -P run."""
    path, words, ped, mp, chrs, files = data
    out = tmp_path / "out.txt"
    opts = ["-t", "i", "-F", str(T1), "-I", str(T2)]
    r = run([EXE, "-A", path, str(out)] + opts + ["-P", files[0][0], "-P", files[1][0]], tmp_path)
    assert "Scanning 2 phenotypes in one batched launch" in r.stdout
    assert not out.exists()
    for j, (f, _) in enumerate(files):
        one = tmp_path / f"single{j}.txt"
        run([EXE, "-A", path, str(one)] + opts + ["-P", f], tmp_path)
        assert open(str(out) + ".P%d" % (j + 1)).read() == open(one).read()


def test_phenotype_files_with_permutations(tmp_path, data):
    """-P f1 -P f2 -p 1-2: four scans in one launch (each file's permutations 1 and 2), each output and each
    perm<k>.P<j>.txt equal to the single run -P fj -p k and its perm<k>.txt."""
    path, words, ped, mp, chrs, files = data
    opts = ["-t", "i", "-F", str(T1), "-I", str(T2)]
    batch = tmp_path / "batch"
    batch.mkdir()
    r = run([EXE, "-A", path, "out.txt"] + opts + ["-P", files[0][0], "-P", files[1][0], "-p", "1-2"], batch)
    assert "Scanning 4 phenotypes in one batched launch" in r.stdout
    for j, (f, _) in enumerate(files):
        for k in (1, 2):
            one = tmp_path / f"one{j}{k}"
            one.mkdir()
            run([EXE, "-A", path, "out.txt"] + opts + ["-P", f, "-p", str(k)], one)
            assert open(batch / ("out.txt.P%d.perm%d" % (j + 1, k))).read() == open(one / "out.txt").read()
            assert open(batch / ("perm%d.P%d.txt" % (k, j + 1))).read() == open(one / ("perm%d.txt" % k)).read()


def test_multi_driver_equals_single_scans(tmp_path, data):
    """epi_perform_scan_multi over the container phenotype, a file's phenotype and a permutation writes what
    one epi_perform_scan per vector writes."""
    import epigpu_amd as E
    from epigpu_amd import scan as S
    path, words, ped, mp, chrs, files = data
    p3 = S.permute_phen(3, ped)
    phen = np.stack([ped["phen"], files[0][1], p3])
    outs = [str(tmp_path / f"m{s}.txt") for s in range(3)]
    S.perform_scan_multi(words, NID, mp, chrs, phen, outs, perms=[0, 0, 3], phenfiles=["", files[0][0], ""],
                         test="i", t1=T1, t2=T2)
    for s in range(3):
        pk = ped.copy()
        pk["phen"] = phen[s]
        one = str(tmp_path / f"s{s}.txt")
        E.perform_scan(words, NID, pk, mp, chrs, one, test="i", t1=T1, t2=T2, perm=[0, 0, 3][s],
                       phenfile=["", files[0][0], ""][s])
        assert open(outs[s]).read() == open(one).read()


def test_dropin_thresholds_used_as_given(tmp_path, data):
    """perform_scan (the drop-in entry) takes UTHRESHOLD1/2 as set, 0 included, as the reference's kernels take
    them (src/main.c:247-249 sets 6.5 / 10.5 only as defaults): with both at 0 it writes what the driver
    writes for t1 = t2 = 0."""
    import epigpu_amd as E
    from epigpu_amd import scan as S
    nid, sizes = 260, [40, 30]
    nsnp = sum(sizes)
    words, y = make_dataset(0x7E0, nid, nsnp, gmiss=0.01, pmiss=0.02)
    y = np.where(np.isnan(y), np.float32(-9.0), y).astype(np.float32)
    ped, mp, chrs = S.make_records(nid, nsnp, sizes, y)
    H = S.epihost()
    glob = {n: ctypes.c_int.in_dll(H, n) for n in ("UTEST", "USAFE", "UPERM", "UDEVICE")}
    t1g, t2g = ctypes.c_float.in_dll(H, "UTHRESHOLD1"), ctypes.c_float.in_dll(H, "UTHRESHOLD2")
    binname = (ctypes.c_char * 100).in_dll(H, "plinkbin")
    uphen = (ctypes.c_char * 1000).in_dll(H, "UPHEN")
    glob["UTEST"].value, glob["USAFE"].value, glob["UPERM"].value, glob["UDEVICE"].value = ord("i"), 0, 0, 0
    t1g.value, t2g.value = 0.0, 0.0
    binname.value, uphen.value = b"data.epigpu", b""
    npack, remain, _ = S.layout(nid)
    w = np.ascontiguousarray(words, dtype=np.int32)
    ped, mp, chrs = (np.ascontiguousarray(x) for x in (ped, mp, chrs))
    out = str(tmp_path / "dropin.txt")
    H.perform_scan.restype = ctypes.c_int
    H.perform_scan.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int] * 5 + [ctypes.c_char_p]
    n = H.perform_scan(w.ctypes.data, chrs.ctypes.data, ped.ctypes.data, mp.ctypes.data, nid, nsnp, npack, remain,
                       len(sizes), out.encode())
    t1g.value, t2g.value = 6.5, 10.5
    ours = str(tmp_path / "ours.txt")
    rep = E.perform_scan(words, nid, ped, mp, chrs, ours, test="i", t1=0.0, t2=0.0)
    assert n == rep["total_hits"] > 1000
    assert open(out).read() == open(ours).read()
