"""ORACLE - test infrastructure only.

The FP64 CPU restatement of the reference's per-pair semantics
(epi_oracle.c, following src/EpiOpenCL.c:45-504 and :1527-1607) and the FP32
reference-shaped CPU scan used as the benchmark's CPU baseline
(epi_refscan32.c). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker; the
product (epigpu_amd/) never does.

Built by oracle/Makefile into oracle/liboracle.so; oracle/_ref/libreadplink.so
is the reference's own readplink.c compiled from /root/reference (codec,
container and the PLINK data mode), used to pin the codec and the ingest.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ORC_HIT_DTYPE = np.dtype([("snp1", "<i4"), ("snp2", "<i4"), ("df1", "<i4"), ("df2", "<i4"),
                          ("F", "<f8"), ("F2", "<f8"), ("block", "<i4"), ("pad", "<i4")])
REF32_DTYPE = np.dtype([("F", "<f4"), ("F2", "<f4"), ("df1", "<i4"), ("df2", "<i4"), ("nfac", "<i4"),
                        ("hit", "<i4")])

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run make -C oracle")
        L = ctypes.CDLL(path)
        vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_float
        L.orc_unpack.argtypes = [vp, i32, i32, vp]
        L.orc_pairs.argtypes = [vp, i32, vp, i32, vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.orc_scan.restype = i64
        L.orc_scan.argtypes = [vp, i32, vp, i32, vp, vp, i32, i32, f32, f32, i32, i32, vp, i64]
        L.ref32_pairs.argtypes = [vp, i32, vp, vp, i64, i32, f32, f32, vp]
        L.ref32_scan_tiles.restype = i64
        L.ref32_scan_tiles.argtypes = [vp, i32, i32, vp, vp, i32, i32, i32, f32, f32, ctypes.POINTER(i64)]
        _lib = L
    return _lib


def ref_readplink():
    """The reference's readplink.c as built by oracle/Makefile, or None."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libreadplink.so")
        if not os.path.exists(path):
            return None
        _ref = ctypes.CDLL(path)
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def unpack(words, nid):
    words = np.ascontiguousarray(words, dtype=np.int32)
    codes = np.empty((words.shape[0], nid), dtype=np.int8)
    lib().orc_unpack(_p(words), nid, words.shape[0], _p(codes))
    return codes


def pairs(codes, phen, pairs_ab, safe=False):
    """FP64 per-pair statistics for an (n, 2) array of (a, b); safe: -9 is a value (the reference's -s 1)."""
    codes = np.ascontiguousarray(codes, dtype=np.int8)
    phen = np.ascontiguousarray(phen, dtype=np.float32)
    pr = np.ascontiguousarray(pairs_ab, dtype=np.int32)
    n = pr.shape[0]
    out = {"counts": np.zeros((n, 9), np.int32), "sums": np.zeros((n, 9), np.float64),
           "F": np.zeros(n, np.float64), "F2": np.zeros(n, np.float64), "df1": np.zeros(n, np.int32),
           "df2": np.zeros(n, np.int32), "nfac": np.zeros(n, np.int32)}
    lib().orc_pairs(_p(codes), codes.shape[1], _p(phen), int(safe), _p(pr), n, _p(out["counts"]), _p(out["sums"]),
                    _p(out["F"]), _p(out["F2"]), _p(out["df1"]), _p(out["df2"]), _p(out["nfac"]))
    return out


def scan(codes, phen, chrstart, chrsize, test="f", t1=6.5, t2=10.5, blk_begin=0, blk_end=1 << 30,
         cap=1 << 22, safe=False):
    """FP64 exhaustive scan; hits sorted by (block, snp1, snp2); safe: the reference's -s 1 semantics."""
    codes = np.ascontiguousarray(codes, dtype=np.int8)
    phen = np.ascontiguousarray(phen, dtype=np.float32)
    cs = np.ascontiguousarray(chrstart, dtype=np.int32)
    cz = np.ascontiguousarray(chrsize, dtype=np.int32)
    hits = np.zeros(cap, dtype=ORC_HIT_DTYPE)
    n = lib().orc_scan(_p(codes), codes.shape[1], _p(phen), int(safe), _p(cs), _p(cz), len(cs), ord(test), t1, t2,
                       blk_begin, blk_end, _p(hits), cap)
    hits = hits[:min(n, cap)]
    hits = hits[np.lexsort((hits["snp2"], hits["snp1"], hits["block"]))]
    return n, hits


def ref32_pairs(words, nid, phen, pairs_ab, test="f", t1=6.5, t2=10.5):
    words = np.ascontiguousarray(words, dtype=np.int32)
    phen = np.ascontiguousarray(phen, dtype=np.float32)
    pr = np.ascontiguousarray(pairs_ab, dtype=np.int32)
    out = np.zeros(pr.shape[0], dtype=REF32_DTYPE)
    lib().ref32_pairs(_p(words), nid, _p(phen), _p(pr), pr.shape[0], ord(test), t1, t2, _p(out))
    return out


def ref32_scan_tiles(words, nid, phen, tiles, tile, test="f", t1=6.5, t2=10.5):
    """Returns (pairs tested, hits) for square tiles given as (a0, b0) rows."""
    words = np.ascontiguousarray(words, dtype=np.int32)
    phen = np.ascontiguousarray(phen, dtype=np.float32)
    tl = np.ascontiguousarray(tiles, dtype=np.int32)
    nh = ctypes.c_int64()
    npairs = lib().ref32_scan_tiles(_p(words), nid, words.shape[0], _p(phen), _p(tl), tl.shape[0], tile, ord(test),
                                    t1, t2, ctypes.byref(nh))
    return npairs, nh.value


def ref_data_mode(pedfile, mapfile, out, clean=False, impute=False, rnorm=False):
    """The reference's data mode -Dr[c][m][n] (src/readplink.c:1330-1435) from its own
    compiled readplink.c: dims, readped, readmap, cleangeno, guessmissing, rnorm, pack,
    createpackedbinary into `out`. Record buffers start zero-filled so that padding
    bytes compare equal, and rand() is seeded with 1 first: the state of the fresh
    process the reference runs in. Returns (nid, nsnp, nchr) or None when the
    reference library is not built. Needs nid, nsnp >= 10 (its progress lines
    divide by n/10)."""
    L = ref_readplink()
    if L is None:
        return None
    c_int, vp = ctypes.c_int, ctypes.c_void_p
    nid, nsnp, nchr = c_int(), c_int(), c_int()
    L.dims(ctypes.byref(nid), ctypes.byref(nsnp), os.fsencode(pedfile))
    assert nid.value >= 10 and nsnp.value >= 10
    dat = ctypes.create_string_buffer(128 * nid.value)
    genmap = ctypes.create_string_buffer(88 * nsnp.value)
    geno = ctypes.create_string_buffer(nid.value * nsnp.value)
    chrstat = ctypes.create_string_buffer(44 * 50)
    L.readped(nid, nsnp, os.fsencode(pedfile), dat, geno, genmap)
    L.readmap(nsnp, genmap, ctypes.byref(nchr), chrstat, os.fsencode(mapfile))
    if clean:
        L.cleangeno(ctypes.byref(nid), ctypes.byref(nsnp), nchr, geno, dat, chrstat, genmap)
        assert nid.value >= 10 and nsnp.value >= 10
    libc = ctypes.CDLL(None)
    libc.srand(1)
    if impute:
        L.guessmissing(nid, nsnp, geno)
    if rnorm:
        L.rnorm(nid, dat)
    npack, remain, gpack = c_int(), c_int(), ctypes.POINTER(c_int)()
    L.pack.argtypes = [c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                       ctypes.POINTER(ctypes.POINTER(c_int)), vp]
    L.pack(nid, nsnp, ctypes.byref(npack), ctypes.byref(remain), ctypes.byref(gpack), geno)
    L.createpackedbinary.argtypes = [c_int, c_int, c_int, c_int, c_int, vp, vp, vp, vp, ctypes.c_char_p]
    L.createpackedbinary(nid, nsnp, npack, remain, nchr, genmap, dat, ctypes.cast(gpack, vp), chrstat,
                         os.fsencode(out))
    libc.free.argtypes = [vp]
    libc.free(ctypes.cast(gpack, vp))
    return nid.value, nsnp.value, nchr.value


def ref_extract(binfile, pedfile, mapfile):
    """The reference's -De (extractfrombinary, src/readplink.c:150-224); False when not built."""
    L = ref_readplink()
    if L is None:
        return False
    L.extractfrombinary(os.fsencode(binfile), os.fsencode(pedfile), os.fsencode(mapfile))
    return True
