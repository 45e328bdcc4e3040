"""The benchmark workloads (synthetic, seeded) as datasets for the scan.

c1  N=500,   M=2000 (1 chr),        MAF U(0.05,0.5), N(0,1),                      8df, p<1e-4
c2  N=4000,  M=20000 (2 x 10000),   MAF U(0.01,0.5), 1% missing calls, N(0,1) 2% NaN, 8df+4df
c3  N=8000,  M=50000 (1 chr),       MAF U(0.05,0.5), (Gamma(2,1)-2)/sqrt2 + 10 planted
                                    pairs (+0.15 SD in one 3x3 cell),             8df+4df
c4  64 independent scans, N ~ U[300,6000], M ~ U[1000,30000] over 1-4 chromosomes,
    MAF U(0.05,0.5), N(0,1) with 3% NaN,                                      8df+4df
c5  N=20000, M=200000 (1 chr),      MAF U(0.01,0.5), 1% missing, 4 phenotypes     8df+4df
Seeds are 0xE9150000 + config number. Thresholds default to the reference's
6.5 / 10.5 (src/main.c:248-249) unless the workload fixes a p-value.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import scan as S

SEED0 = 0xE9150000


@dataclass
class Workload:
    name: str
    nid: int
    chr_sizes: List[int]
    maf: tuple
    gmiss: float
    phen: List[tuple]            # (dist, missing rate) per phenotype
    test: str
    t1: float
    t2: float
    planted: int = 0
    effect: float = 0.15
    seed: int = 0
    words: np.ndarray = field(default=None, repr=False)
    y: np.ndarray = field(default=None, repr=False)
    planted_pairs: np.ndarray = field(default=None, repr=False)

    @property
    def nsnp(self):
        return int(sum(self.chr_sizes))

    @property
    def pairs(self):
        m = self.nsnp
        return m * (m - 1) // 2

    @property
    def chr_starts(self):
        return list(np.cumsum([0] + self.chr_sizes[:-1]))

    def generate(self):
        self.words = S.synth_genotypes(self.seed, self.nid, self.nsnp, self.maf[0], self.maf[1], self.gmiss)
        ys = [S.synth_phenotype(self.seed, r, self.nid, d, m) for r, (d, m) in enumerate(self.phen)]
        self.y = np.stack(ys).astype(np.float32)
        if self.planted:
            self.planted_pairs = S.synth_plant(self.seed, self.words, self.nid, self.y[0], self.planted, self.effect)
        return self


def get(name):
    n = int(name[1:])
    seed = SEED0 + n
    if name == "c1":
        t1 = float(np.float32(S.f_isf(1e-4, 8, 500 - 9)))
        return Workload("c1", 500, [2000], (0.05, 0.5), 0.0, [(S.PHEN_NORMAL, 0.0)], "f", t1, 10.5, seed=seed)
    if name == "c2":
        return Workload("c2", 4000, [10000, 10000], (0.01, 0.5), 0.01, [(S.PHEN_NORMAL, 0.02)], "i", 6.5, 10.5,
                        seed=seed)
    if name == "c3":
        return Workload("c3", 8000, [50000], (0.05, 0.5), 0.0, [(S.PHEN_GAMMA2_STD, 0.0)], "i", 6.5, 10.5,
                        planted=10, seed=seed)
    if name == "c5":
        return Workload("c5", 20000, [200000], (0.01, 0.5), 0.01,
                        [(S.PHEN_NORMAL, 0.0), (S.PHEN_GAMMA2, 0.0), (S.PHEN_T5, 0.0), (S.PHEN_BERN02, 0.0)],
                        "i", 6.5, 10.5, seed=seed)
    raise KeyError(name)


@dataclass
class BatchWorkload:
    """Many independent scans (c4): scan i has its own genotypes, chromosomes and phenotype."""
    name: str
    scans: List[Workload]

    @property
    def pairs(self):
        return int(sum(w.pairs for w in self.scans))

    def generate(self):
        for w in self.scans:
            w.generate()
        return self


def get_batch(name="c4", nscan=64):
    """c4: nscan scans with N ~ U[300,6000], M ~ U[1000,30000] split over 1-4 chromosomes (so both
    intra- and inter-chromosome blocks occur), no missing calls, N(0,1) phenotypes with 3% NaN."""
    if name != "c4":
        raise KeyError(name)
    seed = SEED0 + 4
    rng = np.random.default_rng(seed)
    scans = []
    for i in range(nscan):
        nid = int(rng.integers(300, 6001))
        nsnp = int(rng.integers(1000, 30001))
        nchr = int(rng.integers(1, 5))
        cuts = np.sort(rng.choice(np.arange(1, nsnp), nchr - 1, replace=False)) if nchr > 1 else np.array([], int)
        sizes = np.diff(np.concatenate([[0], cuts, [nsnp]])).astype(int).tolist()
        scans.append(Workload(f"c4.{i}", nid, sizes, (0.05, 0.5), 0.0, [(S.PHEN_NORMAL, 0.03)], "i", 6.5, 10.5,
                              seed=seed * 1000 + i))
    return BatchWorkload(name, scans)
