"""The SPARSE tiles' decomposition of a pair table, on the CPU: the complete-case table over classes {0, 1, NM}
of SNPs a and b rebuilt from (i) the {0, 1} x {0, 1} cells, (ii) per-SNP marginals over all calls and (iii) sums
over the individuals that miss a call in the other SNP, as the kernel combines them,

    X[g][NM] = M_g(a) - sum_{i in miss(b)} I_g(a, i) W(i),   X[NM][h] = M_h(b) - sum_{i in miss(a)} I_h(b, i) W(i),
    sum_cc y^2 = Z_NM(a) - sum_{i in miss(b)} I_NM(a, i) y(i)^2,

then the nine cells by inclusion-exclusion, must give the oracle's counts exactly and its cell sums to rounding,
on data with missing calls in both SNPs of a pair, in one of them, in neither, and missing phenotypes. The same
decomposition with the marginals taken over the pair's complete cases instead of all calls fails. No GPU is
used."""
import numpy as np
import pytest

import oracle as O


def cells_from_x(x):
    c = np.empty(9)
    c[0], c[1], c[3], c[4] = x[0, 0], x[0, 1], x[1, 0], x[1, 1]
    c[2] = x[0, 2] - x[0, 0] - x[0, 1]
    c[5] = x[1, 2] - x[1, 0] - x[1, 1]
    c[6] = x[2, 0] - x[0, 0] - x[1, 0]
    c[7] = x[2, 1] - x[0, 1] - x[1, 1]
    c[8] = x[2, 2] - x[0, 2] - x[1, 2] - x[2, 0] - x[2, 1] + x[0, 0] + x[0, 1] + x[1, 0] + x[1, 1]
    return c


def decompose(codes, y, a, b, marg_over_all=True):
    pres = np.isfinite(y) & (y != -9.0)
    V = pres.astype(float)
    q = np.where(pres, y.astype(np.float64), 0.0)
    ca, cb = codes[a], codes[b]
    miss_a, miss_b = np.flatnonzero(ca == 3), np.flatnonzero(cb == 3)
    ind = lambda c, g: (c != 3) if g == 2 else (c == g)
    # marginals of a SNP: over all its calls (marg_over_all), or wrongly over the pair's complete cases
    both = (ca != 3) & (cb != 3)
    scope = np.ones_like(pres) if marg_over_all else both
    M = lambda c, g, w: (ind(c, g) & scope) @ w
    xc, xs = np.zeros((3, 3)), np.zeros((3, 3))
    for g in range(2):
        for h in range(2):
            m = ind(ca, g) & ind(cb, h)
            xc[g, h], xs[g, h] = m @ V, m @ q
    for g in range(3):  # X[g][NM]: a's class g, b present
        xc[g, 2] = M(ca, g, V) - ind(ca, g)[miss_b] @ V[miss_b]
        xs[g, 2] = M(ca, g, q) - ind(ca, g)[miss_b] @ q[miss_b]
    for h in range(2):  # X[NM][h]
        xc[2, h] = M(cb, h, V) - ind(cb, h)[miss_a] @ V[miss_a]
        xs[2, h] = M(cb, h, q) - ind(cb, h)[miss_a] @ q[miss_a]
    sq = M(ca, 2, q * q) - ind(ca, 2)[miss_b] @ (q * q)[miss_b]
    return cells_from_x(xc), cells_from_x(xs), sq


def dataset(seed, nid=600, nsnp=12):
    rng = np.random.default_rng(seed)
    codes = rng.binomial(2, rng.uniform(0.05, 0.5, nsnp)[:, None], (nsnp, nid)).astype(np.int8)
    codes[rng.random((nsnp, nid)) < 0.03] = 3
    shared = rng.choice(nid, 15, replace=False)  # missing in every SNP: in both SNPs of every pair
    codes[:, shared] = 3
    codes[0][codes[0] == 3] = 0  # SNP 0 has no missing call
    y = rng.standard_normal(nid).astype(np.float32)
    y[rng.random(nid) < 0.03] = np.nan
    y[rng.choice(nid, 5, replace=False)] = -9.0
    return codes, y


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_decomposition_equals_oracle(seed):
    codes, y = dataset(seed)
    nsnp = codes.shape[0]
    pairs = np.array([(a, b) for a in range(nsnp) for b in range(a)], dtype=np.int32)
    ref = O.pairs(codes, y, pairs)
    for k, (a, b) in enumerate(pairs):
        n, s, sq = decompose(codes, y, a, b)
        assert np.array_equal(n, ref["counts"][k].astype(float)), (a, b)
        assert np.allclose(s, ref["sums"][k], rtol=0, atol=1e-9), (a, b)
        ok = (codes[a] != 3) & (codes[b] != 3) & np.isfinite(y) & (y != -9.0)
        assert np.isclose(sq, (y[ok].astype(np.float64) ** 2).sum(), rtol=1e-12)


def test_wrong_marginals_fail():
    """Marginals taken over the pair's complete cases count the missing-in-b individuals twice out."""
    codes, y = dataset(4)
    pairs = np.array([(5, 3)], dtype=np.int32)
    ref = O.pairs(codes, y, pairs)
    n, _, _ = decompose(codes, y, 5, 3, marg_over_all=False)
    assert not np.array_equal(n, ref["counts"][0].astype(float))
