"""Seeded PLINK datasets for the ingest tests, written as text (.ped/.map) and as
PLINK 1 binary (.bed/.bim/.fam, SNP-major). Test infrastructure only."""
import numpy as np

BASES = np.array(list("ACGT"))


def make_dataset(seed, nid=60, nsnp=40, chr_sizes=(15, 15, 10), awkward=True):
    """Genotypes g[nid, nsnp] in {0: A1/A1, 1: A1/A2, 2: A2/A2, -1: missing}.

    awkward adds what the cleaning step removes: SNPs with low call rate, a
    monomorphic SNP, non-positive positions, individuals with low call rate and
    individuals with phenotype -9. Every SNP keeps at least one homozygous call."""
    assert sum(chr_sizes) == nsnp
    rng = np.random.default_rng(seed)
    a = np.stack([rng.choice(4, 2, replace=False) for _ in range(nsnp)])
    a1, a2 = BASES[a[:, 0]], BASES[a[:, 1]]
    p = rng.uniform(0.05, 0.5, nsnp)
    g = (rng.random((nid, nsnp)) < p).astype(int) + (rng.random((nid, nsnp)) < p).astype(int)
    miss = np.full(nsnp, 0.02)
    pd = np.arange(1, nsnp + 1) * 1000
    phen = rng.normal(size=nid).round(4).astype(np.float32)
    if awkward:
        miss[[3, 17, 30]] = 0.3
        g[:, 8] = 0                      # monomorphic
        pd[11], pd[25] = 0, -5
        low = [5, 21, 44]
        for i in low:
            g[i, rng.random(nsnp) < 0.5] = -1
        phen[[2, 33, 50, 51]] = -9
    g[rng.random((nid, nsnp)) < miss] = -1
    for s in range(nsnp):
        if not ((g[:, s] == 0) | (g[:, s] == 2)).any():
            g[rng.integers(nid), s] = 2 * int(rng.integers(2))
    chrs = np.repeat([str(c + 1) for c in range(len(chr_sizes))], chr_sizes)
    return dict(nid=nid, nsnp=nsnp, g=g, a1=a1, a2=a2, phen=phen, pd=pd, chr=chrs,
                gd=rng.uniform(0, 100, nsnp).round(3), sex=rng.integers(1, 3, nid),
                names=[f"rs{1000 + s}" for s in range(nsnp)], seed=seed)


def write_ped_map(ds, ped, mp, het_order="random", zero_first=0.0, sep=" "):
    """het_order: 'random' writes each heterozygote in a random allele order,
    'a1a2' as A1 A2 (what PLINK writes). zero_first: fraction of missing calls
    written '0 A2' instead of '0 0'."""
    rng = np.random.default_rng(ds["seed"] + 1)
    with open(ped, "w") as f:
        for i in range(ds["nid"]):
            ph = "-9" if ds["phen"][i] == -9 else f"{ds['phen'][i]:.4f}"
            row = [f"FAM{i}", f"ID{i}", "0", "0", str(ds["sex"][i]), ph]
            for s in range(ds["nsnp"]):
                x, y = ds["a1"][s], ds["a2"][s]
                v = ds["g"][i, s]
                if v == 0:
                    row += [x, x]
                elif v == 2:
                    row += [y, y]
                elif v == 1:
                    row += [x, y] if het_order == "a1a2" or rng.random() < 0.5 else [y, x]
                else:
                    row += ["0", y] if rng.random() < zero_first else ["0", "0"]
            f.write(sep.join(row) + "\n")
    with open(mp, "w") as f:
        for s in range(ds["nsnp"]):
            f.write(f"{ds['chr'][s]}{sep}{ds['names'][s]}{sep}{ds['gd'][s]}{sep}{ds['pd'][s]}\n")


def write_bed(ds, prefix, a1_override=None):
    """PLINK 1 SNP-major .bed (magic 6c 1b 01; per SNP ceil(N/4) bytes, individual k
    in bits 2(k%4)..2(k%4)+1: 00 A1/A1, 01 missing, 10 A1/A2, 11 A2/A2), .bim and .fam.
    a1_override: {snp: name} to rename A1 (e.g. '0' for a monomorphic SNP)."""
    nid, nsnp = ds["nid"], ds["nsnp"]
    code = {0: 0, 1: 2, 2: 3, -1: 1}
    nb = (nid + 3) // 4
    out = bytearray(b"\x6c\x1b\x01")
    for s in range(nsnp):
        b = bytearray(nb)
        for k in range(nid):
            b[k >> 2] |= code[int(ds["g"][k, s])] << (2 * (k & 3))
        out += b
    open(prefix + ".bed", "wb").write(bytes(out))
    with open(prefix + ".bim", "w") as f:
        for s in range(nsnp):
            x = (a1_override or {}).get(s, ds["a1"][s])
            f.write(f"{ds['chr'][s]}\t{ds['names'][s]}\t{ds['gd'][s]}\t{ds['pd'][s]}\t{x}\t{ds['a2'][s]}\n")
    with open(prefix + ".fam", "w") as f:
        for i in range(nid):
            ph = "-9" if ds["phen"][i] == -9 else f"{ds['phen'][i]:.4f}"
            f.write(f"FAM{i} ID{i} 0 0 {ds['sex'][i]} {ph}\n")
