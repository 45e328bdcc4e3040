"""epi_scan_noalloc, the scan entry point that allocates no device memory, through the C ABI: its hits
(unsorted, FP64 records in a device buffer the caller allocated beforehand) equal those of epi_scan on the same
context, record for record, on three calls in a row, each on a smaller dataset than the one before (the
library's fixed scratch keeps the previous call's tile list and counter), with GENERAL, SPARSE and FAST tiles
in turn; a hit buffer smaller than the hit
count receives a subset with the full count reported; a scan beyond the scratch's tile capacity is refused
before any device work."""
import ctypes

import numpy as np
import pytest

from conftest import make_dataset

pytestmark = pytest.mark.gpu

EPI_ERR_RANGE = -6


class DevBuf:
    """Device memory from the HIP runtime that libepiscan.so itself uses (the test allocates, the entry does not)."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def read(self, dtype, n):
        out = np.zeros(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(n * dtype.itemsize),
                                  2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def sorted_records(h):
    return h[np.lexsort((h["snp2"], h["snp1"]))]


@pytest.mark.parametrize("form", ["mfma", "plain"])
def test_noalloc_equals_scan_on_two_calls(scanner, form):
    from epigpu_amd import scan as S
    cap = 1 << 16
    buf = DevBuf(cap * S.HIT64_DTYPE.itemsize)
    try:
        # GENERAL tiles (1 % missing calls), SPARSE tiles (0.3 %), then a smaller dataset (FAST tiles)
        cases = [(0x40A, 900, [300, 260], 0.01), (0x40D, 700, [200, 180], 0.003), (0x40B, 300, [150], 0.0)]
        for seed, nid, sizes, gmiss in cases:
            nsnp = sum(sizes)
            words, y = make_dataset(seed, nid, nsnp, gmiss=gmiss, pmiss=0.02)
            scanner.set_genotypes(words, nid)
            scanner.set_phenotypes(y)
            scanner.set_chromosomes(list(np.cumsum([0] + sizes[:-1])), sizes)
            kw = dict(test="i", form=form, t1=2.8, t2=3.6)
            g = scanner.scan(**kw)
            kinds = scanner.last_tiles()
            n = scanner.scan_noalloc(buf.ptr.value, cap, **kw)
            assert scanner.last_tiles() == kinds
            if form == "mfma" and gmiss == 0.003:
                assert kinds["sparse"] > 0, kinds
            print(f"noalloc form={form} nid={nid} nsnp={nsnp}: {n} hits (epi_scan {g['nhits']})", flush=True)
            assert n == g["nhits"] > 50 and n <= cap
            h = sorted_records(buf.read(S.HIT64_DTYPE, n))
            order = np.lexsort((g["hits"]["snp2"], g["hits"]["snp1"]))
            ref = g["hits"][order]
            assert (h["snp1"] == ref["snp1"]).all() and (h["snp2"] == ref["snp2"]).all()
            assert (h["df1"] == ref["df1"]).all() and (h["df2"] == ref["df2"]).all()
            assert np.array_equal(h["F"], g["F64"][order]) and np.array_equal(h["F2"], g["F2_64"][order])
            assert (h["tag"] == 0).all()
    finally:
        buf.free()


def test_noalloc_capacity_and_limit(scanner):
    from epigpu_amd import scan as S
    cap = 64
    buf = DevBuf(cap * S.HIT64_DTYPE.itemsize)
    try:
        nid, nsnp = 500, 300
        words, y = make_dataset(0x40C, nid, nsnp, gmiss=0.01, pmiss=0.02)
        scanner.set_genotypes(words, nid)
        scanner.set_phenotypes(y)
        scanner.set_chromosomes([0], [nsnp])
        g = scanner.scan(test="i", form="mfma", t1=2.5, t2=3.0)
        n = scanner.scan_noalloc(buf.ptr.value, cap, test="i", form="mfma", t1=2.5, t2=3.0)
        assert n == g["nhits"] > cap
        h = buf.read(S.HIT64_DTYPE, cap)
        full = set(zip(g["hits"]["snp1"].tolist(), g["hits"]["snp2"].tolist()))
        got = set(zip(h["snp1"].tolist(), h["snp2"].tolist()))
        assert len(got) == cap and got <= full
        # 50,000 SNPs with missing calls need about 1.2 million 32 x 32 tiles: beyond the scratch
        nid, nsnp = 32, 50000
        words, y = make_dataset(0x40D, nid, nsnp, gmiss=0.05, pmiss=0.0)
        scanner.set_genotypes(words, nid)
        scanner.set_phenotypes(y)
        scanner.set_chromosomes([0], [nsnp])
        with pytest.raises(S.EpiError) as e:
            scanner.scan_noalloc(buf.ptr.value, cap, test="i", form="mfma")
        assert e.value.code == EPI_ERR_RANGE and "no-allocation scratch" in str(e.value)
    finally:
        buf.free()
