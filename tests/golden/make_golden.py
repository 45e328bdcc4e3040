"""Writes the golden fixtures under tests/golden/ (run in the build container, needs scipy).

  fdist.json      scipy.stats.f.isf / f.sf values: pins the p -> F threshold helper
  anova.npz       small pair datasets with the 8df F from scipy.stats.f_oneway over the
                  non-empty cells (the reference's F is exactly the one-way ANOVA F over
                  the nfac occupied cells, src/EpiOpenCL.c:231-247)
  interaction.npz pair datasets with proportional cell counts, where the reference's
                  SSI (src/EpiOpenCL.c:463-481) equals the textbook two-way interaction
                  SS_cells - SS_A - SS_B; F2 computed from that decomposition
The reference ships no tests or fixtures of its own, so these library routines
and closed forms are the pins.
"""
import json
import os

import numpy as np
from scipy import stats

HERE = os.path.dirname(os.path.abspath(__file__))


def fdist():
    cases = []
    for p in (1e-2, 1e-3, 1e-4, 1e-6, 1e-8):
        for d1, d2 in ((8, 491), (8, 3991), (4, 3991), (8, 7991), (4, 7991), (8, 19991), (4, 19991), (1, 10),
                       (2, 5), (8, 60), (4, 291)):
            v = float(stats.f.isf(p, d1, d2))
            # scipy's own round trip: where it is loose, scipy's inverse is not a pin
            cases.append({"p": p, "d1": d1, "d2": d2, "isf": v, "sf_at_isf": float(stats.f.sf(v, d1, d2))})
    sf = []
    for F in (0.01, 0.5, 1.0, 2.5, 4.084638, 6.5, 10.5, 30.0):
        for d1, d2 in ((8, 491), (4, 3991), (8, 7991), (3, 17)):
            sf.append({"F": F, "d1": d1, "d2": d2, "sf": float(stats.f.sf(F, d1, d2))})
    with open(os.path.join(HERE, "fdist.json"), "w") as f:
        json.dump({"isf": cases, "sf": sf}, f, indent=1)


def anova():
    rng = np.random.default_rng(20240601)
    A, B, Y, F, NF = [], [], [], [], []
    n = 120
    for case in range(40):
        # genotype frequencies vary so that some cases have empty cells
        pa, pb = rng.uniform(0.03, 0.5, 2)
        ga = rng.binomial(2, pa, n)
        gb = rng.binomial(2, pb, n)
        y = rng.normal(size=n).astype(np.float32)
        if case % 3 == 0:
            y = (y * 3 + 100).astype(np.float32)
        cells = 3 * ga + gb
        groups = [y[cells == c].astype(np.float64) for c in range(9) if (cells == c).any()]
        if len(groups) < 2:
            continue
        f = stats.f_oneway(*groups).statistic
        A.append(ga.astype(np.int8))
        B.append(gb.astype(np.int8))
        Y.append(y)
        F.append(f)
        NF.append(len(groups))
    np.savez(os.path.join(HERE, "anova.npz"), ga=np.array(A), gb=np.array(B), y=np.array(Y), F=np.array(F),
             nfac=np.array(NF))


def interaction():
    """Proportional designs: n_rc = n_r * n_c / N, built as a full cross with replication."""
    rng = np.random.default_rng(7)
    A, B, Y, F2, F = [], [], [], [], []
    for case in range(20):
        ra = rng.integers(1, 4, 3)        # row weights
        cb = rng.integers(1, 4, 3)        # column weights
        rep = int(rng.integers(2, 4))
        ga, gb = [], []
        for r in range(3):
            for c in range(3):
                k = ra[r] * cb[c] * rep
                ga += [r] * k
                gb += [c] * k
        ga, gb = np.array(ga, np.int8), np.array(gb, np.int8)
        n = len(ga)
        eff = rng.normal(size=(3, 3)) * (0.3 if case % 2 else 0.0)
        y = (rng.normal(size=n) + eff[ga, gb] + 0.5 * ga - 0.2 * gb).astype(np.float32)
        perm = rng.permutation(n)
        ga, gb, y = ga[perm], gb[perm], y[perm]
        yd = y.astype(np.float64)
        m = yd.mean()
        cells = 3 * ga + gb
        ss_cells = sum((cells == c).sum() * (yd[cells == c].mean() - m) ** 2 for c in range(9))
        ss_a = sum((ga == r).sum() * (yd[ga == r].mean() - m) ** 2 for r in range(3))
        ss_b = sum((gb == c).sum() * (yd[gb == c].mean() - m) ** 2 for c in range(3))
        ss_w = sum(((yd[cells == c] - yd[cells == c].mean()) ** 2).sum() for c in range(9))
        msw = ss_w / (n - 9)
        A.append(ga)
        B.append(gb)
        Y.append(y)
        F2.append(((ss_cells - ss_a - ss_b) / 4) / msw)
        F.append((ss_cells / 8) / msw)
    lens = {len(a) for a in A}
    width = max(lens)
    pad = lambda arrs, v: np.array([np.concatenate([a, np.full(width - len(a), v, a.dtype)]) for a in arrs])
    # padding individuals carry genotype 3 (missing) so they are excluded
    np.savez(os.path.join(HERE, "interaction.npz"), ga=pad(A, 3), gb=pad(B, 3), y=pad(Y, 0.0), F2=np.array(F2),
             F=np.array(F))


if __name__ == "__main__":
    fdist()
    anova()
    interaction()
    print("fixtures written to", HERE)
