"""The MFMA form's phenotype encoding, on the CPU (epi_phen_encoding is host-only): which values become
outliers, when a phenotype is refused, and that the encoding it returns holds F and F2 within 1e-5 of the FP64
oracle. For the last point the tests restate the kernel's exact integer arithmetic in numpy: q = round((y -
centre) 2^fbits) summed exactly per cell, outliers' values (y - centre) 2^fbits added in FP64, the complete-case
sum of squares from round(q^2 / 2^zbits) plus the outliers' squares; F and F2 then follow in FP64 and are
compared with the oracle's on pairs that both keep and drop the far values (missing calls). A restatement of
the encoding without outliers and with the old 2^23 quantum of q^2 fails the same check, so the check has
teeth. No GPU is used."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err

QMAX = 8355711


def encoding(y):
    from epigpu_amd import scan as S
    return S.phen_encoding(y)


def emulate_F(codes, y, info, pairs, outliers=True, zbits=None):
    """F, F2 of (a, b) pairs from the integer arithmetic of the MFMA form under encoding `info`."""
    pres = np.isfinite(y) & (y != -9.0)
    f, zb = info["fbits"], info["zbits"] if zbits is None else zbits
    dev = y.astype(np.float64) - info["centre"]
    isout = np.zeros(len(y), bool)
    if outliers:
        isout[info["outlier_index"]] = True
    q = np.where(pres & ~isout, np.rint(np.ldexp(dev, f)), 0.0)
    u = np.where(pres & isout, np.ldexp(dev, f), 0.0)
    z = np.rint(q * q / 2.0 ** zb)
    F = np.empty(len(pairs))
    F2 = np.empty(len(pairs))
    for i, (a, b) in enumerate(pairs):
        ok = pres & (codes[a] != 3) & (codes[b] != 3)
        cell = 3 * codes[a][ok].astype(int) + codes[b][ok]
        n = np.bincount(cell, minlength=9).astype(float)
        s = np.bincount(cell, weights=q[ok], minlength=9) + np.bincount(cell, weights=u[ok], minlength=9)
        sq = z[ok].sum() * 2.0 ** zb + (u[ok] ** 2).sum()
        N, tot = n.sum(), s.sum()
        mbar = tot / N
        m = np.where(n > 0, s / np.maximum(n, 1), 0.0)
        nr, nc = n.reshape(3, 3).sum(1), n.reshape(3, 3).sum(0)
        R = np.where(nr > 0, s.reshape(3, 3).sum(1) / np.maximum(nr, 1), 0.0)
        C = np.where(nc > 0, s.reshape(3, 3).sum(0) / np.maximum(nc, 1), 0.0)
        e = (m - mbar)[n > 0]
        ssb = (n[n > 0] * e * e).sum()
        t = (m.reshape(3, 3) - R[:, None] - C[None, :] + mbar).ravel()[n > 0]
        ssi = (n[n > 0] * t * t).sum()
        nfac = (n > 0).sum()
        msw = (sq - N * mbar * mbar - ssb) / (N - nfac)
        F[i] = (ssb / (nfac - 1)) / msw
        F2[i] = (ssi / 4) / msw
    return F, F2


def dataset(seed, nid, nsnp, far):
    """Genotype codes with 1 % missing calls and a phenotype with values far[i] SDs out, each of which misses
    its call on a third of the SNPs."""
    rng = np.random.default_rng(seed)
    maf = rng.uniform(0.05, 0.5, nsnp)
    codes = rng.binomial(2, maf[:, None], (nsnp, nid)).astype(np.int8)
    codes[rng.random((nsnp, nid)) < 0.01] = 3
    y = rng.standard_normal(nid).astype(np.float32)
    idx = rng.choice(nid, len(far), replace=False)
    y[idx] = (np.asarray(far) * np.where(np.arange(len(far)) % 2 == 0, 1.0, -1.0)).astype(np.float32)
    for k in idx:
        codes[rng.random(nsnp) < 1.0 / 3.0, k] = 3
    pairs = np.array([(a, b) for a in range(nsnp) for b in range(a)], dtype=np.int32)
    return codes, y, idx, pairs


def check(codes, y, pairs, info, **kw):
    ref = O.pairs(codes, y, pairs)
    ok = ref["df1"] >= 1
    F, F2 = emulate_F(codes, y, info, pairs, **kw)
    return max(rel_err(F[ok], ref["F"][ok]).max(), rel_err(F2[ok], ref["F2"][ok]).max())


@pytest.mark.parametrize("sd", [30, 100, 1e3, 1e4])
def test_one_far_value_within_tolerance(sd):
    codes, y, idx, pairs = dataset(int(sd), 2000, 26, [sd])
    info = encoding(y)
    assert info["mfma_ok"] == 1
    # an outlier when more than 32 central SDs from the central mean (about 0 here)
    assert info["outliers"] == int(sd > 32 * info["central_sd"])
    if info["outliers"]:
        assert info["outlier_index"] == [int(idx[0])]
    assert info["max_dev"] <= 64 * info["central_sd"]
    assert np.ldexp(info["max_dev"], info["fbits"]) <= QMAX < np.ldexp(info["max_dev"], info["fbits"] + 1)
    assert check(codes, y, pairs, info) <= 1e-5


def test_old_encoding_fails_the_same_check():
    """One value 10^4 SDs out encoded densely (it sets the scale) with q^2 quantized to 2^23: F off by far more
    than 1e-5 on the pairs that drop it."""
    codes, y, idx, pairs = dataset(7, 2000, 26, [1e4])
    info = dict(encoding(y))
    pres = np.isfinite(y)
    info["centre"] = float(y[pres].astype(np.float64).mean())
    maxabs = np.abs(y[pres] - info["centre"]).max()
    info["fbits"] = int(np.floor(np.log2(QMAX / maxabs)))
    assert check(codes, y, pairs, info, outliers=False, zbits=23) > 1e-3


def test_many_far_values_and_refusal():
    """Up to 32 far values are carried exactly; with 40 the rest would set a coarse scale and the MFMA form
    refuses the phenotype (mfma_ok = 0)."""
    codes, y, idx, pairs = dataset(11, 2000, 20, list(np.geomspace(40, 1e5, 20)))
    info = encoding(y)
    assert info["mfma_ok"] == 1 and info["outliers"] == 20 and sorted(info["outlier_index"]) == sorted(idx.tolist())
    assert check(codes, y, pairs, info) <= 1e-5
    _, y40, _, _ = dataset(12, 2000, 4, [1e4] * 40)
    info = encoding(y40)
    assert info["mfma_ok"] == 0 and info["outliers"] == 32 and info["max_dev"] > 64 * info["central_sd"]


@pytest.mark.parametrize("trait", ["lognormal1", "lognormal2", "gamma", "t3", "cauchy", "binary0.002"])
def test_raw_traits_within_tolerance(trait):
    rng = np.random.default_rng(["lognormal1", "lognormal2", "gamma", "t3", "cauchy", "binary0.002"].index(trait))
    codes, _, _, pairs = dataset(21, 3000, 20, [])
    n = codes.shape[1]
    y = {"lognormal1": lambda: np.exp(rng.standard_normal(n)), "lognormal2": lambda: np.exp(2 * rng.standard_normal(n)),
         "gamma": lambda: rng.gamma(2.0, 1.0, n), "t3": lambda: rng.standard_t(3, n),
         "cauchy": lambda: rng.standard_cauchy(n),
         "binary0.002": lambda: (rng.random(n) < 0.002).astype(float)}[trait]().astype(np.float32)
    y[rng.random(n) < 0.02] = np.nan
    info = encoding(y)
    assert info["present"] == int(np.isfinite(y).sum())
    if info["mfma_ok"]:
        assert check(codes, y, pairs, info) <= 1e-5


def test_encoding_errors_and_sentinel():
    from epigpu_amd import scan as S
    y = np.ones(50, np.float32)
    y[3] = np.inf
    with pytest.raises(S.EpiError) as e:
        encoding(y)
    assert e.value.code == -6
    with pytest.raises(S.EpiError):
        encoding(np.full(40, -9.0, np.float32))
    y = np.random.default_rng(0).standard_normal(300).astype(np.float32)
    y[:5] = -9.0
    assert encoding(y)["present"] == 295
    assert S.phen_encoding(y, nine_is_value=True)["present"] == 300
