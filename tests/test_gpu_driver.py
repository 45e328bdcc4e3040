"""GPU tests of the host driver: perform_scan's output file against the FP64 oracle,
kill-and-resume, the CLI, and the reference's own main() linked against libepihost."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import make_dataset

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SIZES = [300, 220, 180]
NID = 700
T1, T2 = 3.2, 4.2


@pytest.fixture(scope="module")
def dataset():
    import epigpu_amd as E
    from epigpu_amd import scan as S
    nsnp = sum(SIZES)
    words, y = make_dataset(31, NID, nsnp, maf=(0.02, 0.5), gmiss=0.01, pmiss=0.02)
    # the container stores phenotypes in ped records; -9 is the reference's missing code there
    y = np.where(np.isnan(y), np.float32(-9.0), y).astype(np.float32)
    ped, mp, chrs = S.make_records(NID, nsnp, SIZES, y)
    return words, y, ped, mp, chrs


def parse_output(path):
    blocks, rows, header = [], {}, []
    cur = None
    for line in open(path):
        m = re.match(r"# (\d+) x (\d+) : (\d+) interactions", line)
        if m:
            cur = (int(m.group(1)), int(m.group(2)))
            blocks.append((cur, int(m.group(3))))
            rows[cur] = []
        elif cur is not None:
            f = line.rstrip("\n").split("\t")
            rows[cur].append((int(f[0]) - 1, int(f[3]) - 1, int(f[6]), int(f[7]), float(f[8]), float(f[9])))
        else:
            header.append(line)
    return header, blocks, rows


def test_perform_scan_output_matches_oracle(tmp_path, dataset):
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    out = str(tmp_path / "out.txt")
    rep = E.perform_scan(words, NID, ped, mp, chrs, out, test="i", t1=T1, t2=T2)
    header, blocks, rows = parse_output(out)
    assert header[0] == "Datafile: data.epigpu \n"
    assert header[5] == "Test: interactions + full\n"
    assert [b for b, _ in blocks] == [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    codes = O.unpack(words, NID)
    starts = [0, 300, 520]
    n_ref, ref = O.scan(codes, y, starts, SIZES, "i", T1, T2)
    assert rep["total_hits"] == sum(k for _, k in blocks) and rep["blocks_done"] == 6
    for blk, ((i, j), k) in enumerate(blocks):
        r = ref[ref["block"] == blk]
        got = rows[(i, j)]
        assert len(got) == k
        gs = {(a, b) for a, b, *_ in got}
        rs = set(zip(r["snp1"].tolist(), r["snp2"].tolist()))
        for pr in gs ^ rs:  # only threshold ties may differ
            x = r[(r["snp1"] == pr[0]) & (r["snp2"] == pr[1])]
            assert len(x) == 0 or min(abs(x["F"][0] - T1) / T1, abs(x["F2"][0] - T2) / T2) < 1e-5
        order = [(a, b) for a, b, *_ in got]
        assert order == sorted(order)
        rd = {(a, b): (d1, d2, F, F2) for a, b, d1, d2, F, F2 in zip(r["snp1"], r["snp2"], r["df1"], r["df2"],
                                                                       r["F"], r["F2"])}
        for a, b, d1, d2, F, F2 in got:
            if (a, b) in rd:
                e = rd[(a, b)]
                assert (d1, d2) == (e[0], e[1])
                assert abs(F - e[2]) <= 1e-5 * max(1.0, e[2]) + 1e-6
                assert abs(F2 - e[3]) <= 1e-5 * max(1.0, e[3]) + 1e-6


def test_resume_after_interrupted_scan(tmp_path, dataset):
    """Kill after block 3 (mid-append), rerun: the file equals an uninterrupted run."""
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    full = str(tmp_path / "full.txt")
    E.perform_scan(words, NID, ped, mp, chrs, full, test="i", t1=T1, t2=T2)
    text = open(full).read()
    cut = text.index("# 2 x 0 :")  # blocks (0,0) (1,0) (1,1) complete
    part = str(tmp_path / "part.txt")
    with open(part, "w") as f:
        f.write(text[:cut] + "# 2 x 0 : 5 interactions\n1\trs")  # interrupted append
    rep = E.perform_scan(words, NID, ped, mp, chrs, part, test="i", t1=T1, t2=T2)
    assert rep["first_block"] == 3 and rep["blocks_done"] == 3
    assert open(part).read() == text
    again = E.perform_scan(words, NID, ped, mp, chrs, part, test="i", t1=T1, t2=T2)
    assert again["blocks_done"] == 0 and open(part).read() == text


def test_resume_refuses_other_test(tmp_path, dataset):
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    out = str(tmp_path / "out.txt")
    E.perform_scan(words, NID, ped, mp, chrs, out, test="i", t1=T1, t2=T2)
    with pytest.raises(E.EpiError):
        E.perform_scan(words, NID, ped, mp, chrs, out, test="f", t1=T1, t2=T2)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_cli_writes_same_file(tmp_path, dataset, form):
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    data = str(tmp_path / "data.epigpu")
    E.write_container(data, words, NID, ped, mp, chrs)
    ref_out = str(tmp_path / "ref.txt")
    E.perform_scan(words, NID, ped, mp, chrs, ref_out, test="i", t1=T1, t2=T2)
    cli_out = str(tmp_path / "cli.txt")
    r = subprocess.run([os.path.join(ROOT, "epigpu_amd", "lib", "episcan"), "-A", data, cli_out, "-t", "i", "-F",
                        str(T1), "-I", str(T2), "-m", form], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b = parse_output(ref_out), parse_output(cli_out)
    assert a[0] == b[0] and a[1] == b[1]
    for k in a[2]:
        sa = [(x[0], x[1], x[2], x[3]) for x in a[2][k]]
        sb = [(x[0], x[1], x[2], x[3]) for x in b[2][k]]
        assert sa == sb


def test_reference_main_dropin(tmp_path, dataset):
    """The reference's own analysis-mode main() (src/main.c) linked against libepihost.so."""
    exe = os.path.join(ROOT, "oracle", "_ref", "epiGPU-dropin")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built (needs the reference tree at build time)")
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    data = str(tmp_path / "data.epigpu")
    E.write_container(data, words, NID, ped, mp, chrs)
    out = str(tmp_path / "dropin.txt")
    r = subprocess.run([exe, "-A", data, out, "-t", "i", "-F", str(T1), "-I", str(T2)], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ours = str(tmp_path / "ours.txt")
    E.perform_scan(words, NID, ped, mp, chrs, ours, test="i", t1=T1, t2=T2)
    assert open(out).read() == open(ours).read()


def test_batched_permutations_equal_single_runs(tmp_path, dataset):
    """Permutations 1..3 in one batched launch write the same files as three single runs of permutation k
    (phenotypes shuffled as the reference's permute() with srand(k))."""
    import epigpu_amd as E
    from epigpu_amd import scan as S
    words, y, ped, mp, chrs = dataset
    fmt = str(tmp_path / "batch.perm%d")
    rep = S.perform_scan_perms(words, NID, ped, mp, chrs, fmt, 1, 3, test="i", t1=T1, t2=T2)
    assert rep["blocks_done"] == 3 * 6
    for k in (1, 2, 3):
        pk = ped.copy()
        pk["phen"] = S.permute_phen(k, ped)
        one = str(tmp_path / f"single{k}.txt")
        E.perform_scan(words, NID, pk, mp, chrs, one, test="i", t1=T1, t2=T2, perm=k)
        assert open(fmt % k).read() == open(one).read()
        assert "Permutation: %d\n" % k in open(one).read()


def test_cli_permutation_range(tmp_path, dataset):
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    data = str(tmp_path / "data.epigpu")
    E.write_container(data, words, NID, ped, mp, chrs)
    out = str(tmp_path / "cli.txt")
    exe = os.path.join(ROOT, "epigpu_amd", "lib", "episcan")
    r = subprocess.run([exe, "-A", data, out, "-t", "i", "-F", str(T1), "-I", str(T2), "-p", "2-3"],
                       capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    one = tmp_path / "one"
    one.mkdir()
    single = str(one / "single.txt")
    r = subprocess.run([exe, "-A", data, single, "-t", "i", "-F", str(T1), "-I", str(T2), "-p", "3"],
                       capture_output=True, text=True, timeout=120, cwd=str(one))
    assert r.returncode == 0, r.stderr
    assert open(out + ".perm3").read() == open(single).read()
    assert os.path.exists(out + ".perm2") and not os.path.exists(out)
    # the permute() side files (src/main.c:349-356), one per permutation, in the working directory
    assert open(tmp_path / "perm3.txt").read() == open(one / "perm3.txt").read()
    assert (tmp_path / "perm2.txt").exists()


def test_safe_mode_cli_and_reference_main(tmp_path, dataset):
    """-s 1 through the CLI, perform_scan(safe=True) and the reference's own main() write the same file."""
    import epigpu_amd as E
    words, y, ped, mp, chrs = dataset
    data = str(tmp_path / "data.epigpu")
    E.write_container(data, words, NID, ped, mp, chrs)
    ours = str(tmp_path / "ours.txt")
    E.perform_scan(words, NID, ped, mp, chrs, ours, test="i", t1=T1, t2=T2, safe=True)
    normal = str(tmp_path / "normal.txt")
    E.perform_scan(words, NID, ped, mp, chrs, normal, test="i", t1=T1, t2=T2)
    assert open(ours).read() != open(normal).read()
    cli = str(tmp_path / "cli.txt")
    r = subprocess.run([os.path.join(ROOT, "epigpu_amd", "lib", "episcan"), "-A", data, cli, "-t", "i", "-F", str(T1),
                        "-I", str(T2), "-s", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert open(cli).read() == open(ours).read()
    exe = os.path.join(ROOT, "oracle", "_ref", "epiGPU-dropin")
    if os.path.exists(exe):
        ref = str(tmp_path / "dropin.txt")
        r = subprocess.run([exe, "-A", data, ref, "-t", "i", "-F", str(T1), "-I", str(T2), "-s", "1"],
                           capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert open(ref).read() == open(ours).read()


def test_plink_ingest_then_scan(tmp_path):
    """PLINK text -> container (-Drcm) -> scan (-A -t i): the hits are the oracle's on the
    container's genotypes and phenotypes (33 SNPs after cleaning, low thresholds)."""
    import epigpu_amd as E
    exe = os.path.join(ROOT, "epigpu_amd", "lib", "episcan")
    gold = os.path.join(ROOT, "tests", "golden")
    data = str(tmp_path / "p.epigpu")
    r = subprocess.run([exe, "-Drcm", os.path.join(gold, "plink_small.ped"), os.path.join(gold, "plink_small.map"),
                        data], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    c = E.read_container(data)
    out = str(tmp_path / "out.txt")
    r = subprocess.run([exe, "-A", data, out, "-t", "i", "-F", "1.5", "-I", "1.5"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    header, blocks, rows = parse_output(out)
    sizes = c["chr"]["chrsize"].tolist()
    starts = c["chr"]["chrstart"].tolist()
    n_ref, ref = O.scan(O.unpack(c["geno"], c["nid"]), c["ped"]["phen"], starts, sizes, "i", 1.5, 1.5)
    got = {(a, b) for rs in rows.values() for a, b, *_ in rs}
    want = set(zip(ref["snp1"].tolist(), ref["snp2"].tolist()))
    assert n_ref > 10
    for pr in got ^ want:  # only threshold ties may differ
        x = ref[(ref["snp1"] == pr[0]) & (ref["snp2"] == pr[1])]
        assert len(x) == 0 or min(abs(x["F"][0] - 1.5), abs(x["F2"][0] - 1.5)) < 1e-5 * 1.5
    assert len(got ^ want) <= 2


def test_empty_chromosome_blocks(tmp_path):
    """A chromosome left with no SNPs (as cleaning can leave one) keeps its blocks in the output, each with
    0 interactions, and the other blocks' hits are the oracle's."""
    import epigpu_amd as E
    from epigpu_amd import scan as S
    nid, sizes = 400, [60, 0, 45]
    nsnp = sum(sizes)
    words, y = make_dataset(77, nid, nsnp, maf=(0.05, 0.5), gmiss=0.01, pmiss=0.02)
    y = np.where(np.isnan(y), np.float32(-9.0), y).astype(np.float32)
    ped, mp, chrs = S.make_records(nid, nsnp, sizes, y)
    out = str(tmp_path / "out.txt")
    rep = E.perform_scan(words, nid, ped, mp, chrs, out, test="i", t1=T1, t2=T2)
    header, blocks, rows = parse_output(out)
    assert [b for b, _ in blocks] == [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    assert dict(blocks)[(1, 0)] == dict(blocks)[(1, 1)] == dict(blocks)[(2, 1)] == 0
    n_ref, ref = O.scan(O.unpack(words, nid), y, [0, 60, 60], sizes, "i", T1, T2)
    got = {(a, b) for rs in rows.values() for a, b, *_ in rs}
    want = set(zip(ref["snp1"].tolist(), ref["snp2"].tolist()))
    assert n_ref > 10 and rep["total_hits"] == len(got)
    for pr in got ^ want:  # only threshold ties may differ
        x = ref[(ref["snp1"] == pr[0]) & (ref["snp2"] == pr[1])]
        assert len(x) == 0 or min(abs(x["F"][0] - T1) / T1, abs(x["F2"][0] - T2) / T2) < 1e-5
