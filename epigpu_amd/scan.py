"""ctypes bindings of the native libraries (plumbing for tests and the benchmark).

The product is libepiscan.so (episcan.h) and libepihost.so (epihost.h); this
module only moves numpy arrays across their C ABI. Loading fails loudly when a
library has not been built: there is no Python fallback for any computation.
"""
import ctypes
import weakref
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

FORM_PLAIN = 0
FORM_MFMA = 1
FORMS = {"plain": FORM_PLAIN, "mfma": FORM_MFMA}
PHEN_NINE_IS_VALUE = 1

PHEN_NORMAL, PHEN_GAMMA2, PHEN_GAMMA2_STD, PHEN_T5, PHEN_BERN02 = range(5)


class EpiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class Hit(ctypes.Structure):
    """Reference `table` record (src/id4117.h:14-21)."""
    _fields_ = [("snp1", ctypes.c_int32), ("snp2", ctypes.c_int32), ("df1", ctypes.c_int32),
                ("df2", ctypes.c_int32), ("F", ctypes.c_float), ("F2", ctypes.c_float)]


HIT_DTYPE = np.dtype([("snp1", "<i4"), ("snp2", "<i4"), ("df1", "<i4"), ("df2", "<i4"),
                      ("F", "<f4"), ("F2", "<f4")])


class PhenInfo(ctypes.Structure):
    """epi_phen_info: how the MFMA form encodes a phenotype vector."""
    _fields_ = [("present", ctypes.c_int32), ("outliers", ctypes.c_int32), ("mfma_ok", ctypes.c_int32),
                ("fbits", ctypes.c_int32), ("zbits", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("centre", ctypes.c_double), ("central_sd", ctypes.c_double), ("max_dev", ctypes.c_double),
                ("outlier_index", ctypes.c_int32 * 32)]


class ScanParams(ctypes.Structure):
    _fields_ = [("test", ctypes.c_int32), ("form", ctypes.c_int32), ("t1", ctypes.c_float),
                ("t2", ctypes.c_float), ("phen", ctypes.c_int32), ("blk_begin", ctypes.c_int32),
                ("blk_end", ctypes.c_int32), ("safe", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class Schedule(ctypes.Structure):
    _fields_ = [("fast_rows", ctypes.c_int32), ("super_a", ctypes.c_int32), ("super_b", ctypes.c_int32),
                ("queues", ctypes.c_int32), ("order", ctypes.c_int32), ("dense_missing", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 2)]

    def as_dict(self):
        return {k: getattr(self, k) for k in ("fast_rows", "super_a", "super_b", "queues", "order", "dense_missing")}


class TuneEntry(ctypes.Structure):
    _fields_ = [("sched", Schedule), ("ms", ctypes.c_double)]


TUNE_ROWS, TUNE_SUPER, TUNE_QUEUES, TUNE_ORDER, TUNE_MISSING, TUNE_ALL = 1, 2, 4, 8, 16, 31


class RunOpts(ctypes.Structure):
    _fields_ = [("test", ctypes.c_int), ("t1", ctypes.c_float), ("t2", ctypes.c_float),
                ("device", ctypes.c_int), ("form", ctypes.c_int), ("perm", ctypes.c_int),
                ("datafile", ctypes.c_char_p), ("phenfile", ctypes.c_char_p), ("quiet", ctypes.c_int),
                ("safe", ctypes.c_int)]


class RunReport(ctypes.Structure):
    _fields_ = [("total_hits", ctypes.c_int64), ("new_hits", ctypes.c_int64), ("blocks_done", ctypes.c_int),
                ("first_block", ctypes.c_int), ("overflow", ctypes.c_int), ("kernel_s", ctypes.c_double),
                ("wall_s", ctypes.c_double), ("pairs", ctypes.c_int64)]


class ResumeState(ctypes.Structure):
    _fields_ = [("exists", ctypes.c_int), ("chr1", ctypes.c_int), ("chr2", ctypes.c_int),
                ("tothits", ctypes.c_int64), ("blocks", ctypes.c_int)]


class Plink(ctypes.Structure):
    _fields_ = [("nid", ctypes.c_int), ("nsnp", ctypes.c_int), ("nchr", ctypes.c_int), ("dat", ctypes.c_void_p),
                ("genmap", ctypes.c_void_p), ("chrstat", ctypes.c_void_p), ("codes", ctypes.c_void_p),
                ("snpchr", ctypes.c_void_p)]


class PlinkReport(ctypes.Structure):
    _fields_ = [("snps_in", ctypes.c_int), ("snps_kept", ctypes.c_int), ("ids_in", ctypes.c_int),
                ("ids_kept", ctypes.c_int), ("imputed", ctypes.c_int64)]


PLINK_CLEAN, PLINK_IMPUTE, PLINK_RNORM = 1, 2, 4


class Container(ctypes.Structure):
    _fields_ = [("nid", ctypes.c_int), ("nsnp", ctypes.c_int), ("npack", ctypes.c_int), ("remain", ctypes.c_int),
                ("nchr", ctypes.c_int), ("dat", ctypes.c_void_p), ("genmap", ctypes.c_void_p),
                ("chrstat", ctypes.c_void_p), ("geno", ctypes.c_void_p)]


# Record layouts of the epiGPU container (src/readplink.h:17-39, x86-64 gcc).
HIT64_DTYPE = np.dtype([("snp1", "<i4"), ("snp2", "<i4"), ("df1", "<i4"), ("df2", "<i4"), ("F", "<f8"),
                        ("F2", "<f8"), ("tag", "<i4"), ("pad", "<i4")])
PED_DTYPE = np.dtype({"names": ["family", "id", "paternal", "maternal", "sex", "phen"],
                      "formats": ["S30", "S30", "S30", "S30", "S1", "<f4"],
                      "offsets": [0, 30, 60, 90, 120, 124], "itemsize": 128})
MAP_DTYPE = np.dtype({"names": ["chr", "snpname", "gd", "pd", "allelename"],
                      "formats": ["S30", "S30", "<f8", "<i8", "S2"],
                      "offsets": [0, 30, 64, 72, 80], "itemsize": 88})
CHR_DTYPE = np.dtype({"names": ["chrid", "chrname", "chrsize", "chrstart"],
                      "formats": ["<i4", "S30", "<i4", "<i4"],
                      "offsets": [0, 4, 36, 40], "itemsize": 44})

_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(LIBDIR, name)
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run __graft_entry__.build()")
        _libs[name] = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        _declare(name, _libs[name])
    return _libs[name]


def _p(t):
    return ctypes.POINTER(t)


def _declare(name, L):
    i32, i64, f32, f64, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
    if name == "libepiscan.so":
        sig = {
            "epi_open": (ctypes.c_int, [ctypes.c_int, _p(vp)]),
            "epi_close": (None, [vp]),
            "epi_strerror": (ctypes.c_char_p, [ctypes.c_int]),
            "epi_last_error": (ctypes.c_char_p, [vp]),
            "epi_device_name": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t]),
            "epi_set_genotypes": (ctypes.c_int, [vp, vp, i32, i32]),
            "epi_set_phenotypes": (ctypes.c_int, [vp, vp, i32, i32]),
            "epi_set_phenotypes_ex": (ctypes.c_int, [vp, vp, i32, i32, i32]),
            "epi_set_chromosomes": (ctypes.c_int, [vp, vp, vp, i32]),
            "epi_scan_params_default": (None, [_p(ScanParams)]),
            "epi_scan": (ctypes.c_int, [vp, _p(ScanParams)]),
            "epi_phen_encoding": (ctypes.c_int, [vp, i32, i32, _p(PhenInfo)]),
            "epi_scan_noalloc": (ctypes.c_int, [vp, _p(ScanParams), vp, i64, _p(i64)]),
            "epi_hit_count": (ctypes.c_int, [vp, _p(i64), _p(i32)]),
            "epi_get_hits": (ctypes.c_int, [vp, vp, vp, vp, vp, i64, _p(i64)]),
            "epi_set_hit_capacity": (ctypes.c_int, [vp, i64]),
            "epi_last_scan_stats": (ctypes.c_int, [vp, _p(f64), _p(i64), _p(i32)]),
            "epi_last_tiles": (ctypes.c_int, [vp, _p(i64), _p(i64), _p(i64)]),
            "epi_pair_dump": (ctypes.c_int, [vp, _p(ScanParams), i32, i32, i32, i32, vp, vp, vp, vp]),
            "epi_batch_create": (ctypes.c_int, [vp, _p(vp)]),
            "epi_batch_destroy": (None, [vp]),
            "epi_batch_add_dataset": (ctypes.c_int, [vp, vp, i32, i32, vp, vp, i32, _p(i32)]),
            "epi_batch_add_phenotypes": (ctypes.c_int, [vp, i32, vp, i32, _p(i32)]),
            "epi_batch_add_phenotypes_ex": (ctypes.c_int, [vp, i32, vp, i32, i32, _p(i32)]),
            "epi_batch_set_hit_capacity": (ctypes.c_int, [vp, i64]),
            "epi_batch_scan": (ctypes.c_int, [vp, _p(ScanParams)]),
            "epi_batch_hit_count": (ctypes.c_int, [vp, i32, _p(i64), _p(i32)]),
            "epi_batch_get_hits": (ctypes.c_int, [vp, i32, vp, vp, vp, vp, i64, _p(i64)]),
            "epi_batch_stats": (ctypes.c_int, [vp, _p(f64), _p(i64), _p(i64), _p(i32), _p(i32)]),
            "epi_set_max_tiles": (ctypes.c_int, [vp, i64]),
            "epi_set_schedule": (ctypes.c_int, [vp, i32, _p(Schedule)]),
            "epi_autotune": (ctypes.c_int, [vp, _p(ScanParams), i32, i32, _p(Schedule), _p(f64), _p(TuneEntry), i32,
                                            _p(i32)]),
            "epi_get_schedule": (ctypes.c_int, [vp, i32, _p(Schedule)]),
            "epi_batch_set_schedule": (ctypes.c_int, [vp, i32, _p(Schedule)]),
            "epi_batch_get_schedule": (ctypes.c_int, [vp, i32, _p(Schedule)]),
            "epi_batch_autotune": (ctypes.c_int, [vp, _p(ScanParams), i32, i32, _p(Schedule), _p(f64),
                                                  _p(TuneEntry), i32, _p(i32)]),
        }
    else:
        sig = {
            "epi_layout": (None, [ctypes.c_int, _p(ctypes.c_int), _p(ctypes.c_int), _p(ctypes.c_int)]),
            "epi_pack": (None, [vp, ctypes.c_int, ctypes.c_int, vp]),
            "epi_unpack": (None, [vp, ctypes.c_int, ctypes.c_int, vp]),
            "epi_synth_genotypes": (None, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, f64, f64, f64, vp]),
            "epi_synth_phenotype": (None, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64, vp]),
            "epi_synth_plant": (None, [ctypes.c_uint64, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, f32, vp]),
            "epi_betai": (f64, [f64, f64, f64]),
            "epi_f_sf": (f64, [f64, f64, f64]),
            "epi_f_isf": (f64, [f64, f64, f64]),
            "epi_read_container": (ctypes.c_int, [ctypes.c_char_p, _p(Container)]),
            "epi_write_container": (ctypes.c_int, [ctypes.c_char_p, _p(Container)]),
            "epi_container_free": (None, [_p(Container)]),
            "epih_strerror": (ctypes.c_char_p, [ctypes.c_int]),
            "epi_run_opts_default": (None, [_p(RunOpts)]),
            "epi_perform_scan": (ctypes.c_int, [_p(RunOpts), vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                                _p(RunReport)]),
            "epi_permute_phen": (None, [ctypes.c_uint, ctypes.c_int, vp, vp]),
            "epi_permute_order": (None, [ctypes.c_uint, ctypes.c_int, vp]),
            "epi_write_perm_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, vp, vp, vp]),
            "epi_perform_scan_multi": (ctypes.c_int, [_p(RunOpts), vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                                      vp, vp, vp, _p(RunReport)]),
            "epi_perform_scan_perms": (ctypes.c_int, [_p(RunOpts), vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_int, ctypes.c_char_p, _p(RunReport)]),
            "epi_resume": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, f32, f32, _p(ResumeState)]),
            "epi_read_ped_map": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _p(Plink)]),
            "epi_read_bed": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _p(Plink)]),
            "epi_plink_process": (ctypes.c_int, [_p(Plink), ctypes.c_int, ctypes.c_uint, _p(PlinkReport)]),
            "epi_plink_to_container": (ctypes.c_int, [_p(Plink), _p(Container)]),
            "epi_plink_free": (None, [_p(Plink)]),
            "epi_write_ped_map": (ctypes.c_int, [_p(Container), ctypes.c_char_p, ctypes.c_char_p]),
            "epi_plink_error": (ctypes.c_char_p, []),
            "epi_format_row": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                              ctypes.c_char_p, ctypes.c_int, ctypes.c_int, f32, f32]),
        }
    for fn, (res, args) in sig.items():
        f = getattr(L, fn)
        f.restype = res
        f.argtypes = args


def episcan():
    return _load("libepiscan.so")


def epihost():
    episcan()
    return _load("libepihost.so")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ------------------------------------------------------------- host side --

def layout(nid):
    npack, remain = nid // 16, nid % 16
    return npack, remain, npack + (remain > 0)


def pack(codes):
    """codes: (nsnp, nid) int8 in {0,1,2,3} -> (nsnp, npack1) int32 words (layout A1)."""
    codes = np.ascontiguousarray(codes, dtype=np.int8)
    nsnp, nid = codes.shape
    words = np.empty((nsnp, layout(nid)[2]), dtype=np.int32)
    epihost().epi_pack(_ptr(codes), nid, nsnp, _ptr(words))
    return words


def unpack(words, nid):
    words = np.ascontiguousarray(words, dtype=np.int32)
    nsnp = words.shape[0]
    codes = np.empty((nsnp, nid), dtype=np.int8)
    epihost().epi_unpack(_ptr(words), nid, nsnp, _ptr(codes))
    return codes


def synth_genotypes(seed, nid, nsnp, maf_lo=0.05, maf_hi=0.5, miss=0.0):
    words = np.empty((nsnp, layout(nid)[2]), dtype=np.int32)
    epihost().epi_synth_genotypes(seed, nid, nsnp, maf_lo, maf_hi, miss, _ptr(words))
    return words


def synth_phenotype(seed, r, nid, dist=PHEN_NORMAL, miss=0.0):
    y = np.empty(nid, dtype=np.float32)
    epihost().epi_synth_phenotype(seed, r, nid, dist, miss, _ptr(y))
    return y


def synth_plant(seed, words, nid, phen, npairs, effect):
    words = np.ascontiguousarray(words, dtype=np.int32)
    pairs = np.zeros((npairs, 4), dtype=np.int32)
    epihost().epi_synth_plant(seed, _ptr(words), nid, words.shape[0], _ptr(phen), npairs, effect, _ptr(pairs))
    return pairs


def f_sf(F, d1, d2):
    return epihost().epi_f_sf(F, d1, d2)


def f_isf(p, d1, d2):
    return epihost().epi_f_isf(p, d1, d2)


def betai(a, b, x):
    return epihost().epi_betai(a, b, x)


def phen_encoding(phen, nine_is_value=False):
    """How the MFMA form encodes one phenotype vector (epi_phen_encoding; host only)."""
    y = np.ascontiguousarray(phen, dtype=np.float32)
    info = PhenInfo()
    rc = episcan().epi_phen_encoding(_ptr(y), len(y), PHEN_NINE_IS_VALUE if nine_is_value else 0,
                                     ctypes.byref(info))
    if rc:
        raise EpiError(rc, "epi_phen_encoding: " + episcan().epi_strerror(rc).decode())
    out = {f: getattr(info, f) for f, _ in PhenInfo._fields_ if f not in ("pad", "outlier_index")}
    out["outlier_index"] = list(info.outlier_index)[:info.outliers]
    return out


def read_container(path):
    c = Container()
    rc = epihost().epi_read_container(os.fsencode(path), ctypes.byref(c))
    if rc:
        raise EpiError(rc, epihost().epih_strerror(rc).decode())
    try:
        npack1 = c.npack + (c.remain > 0)
        def arr(ptr, dtype, n):
            buf = (ctypes.c_char * (dtype.itemsize * n)).from_address(ptr)
            return np.frombuffer(bytearray(buf), dtype=dtype)
        out = {
            "nid": c.nid, "nsnp": c.nsnp, "npack": c.npack, "remain": c.remain, "nchr": c.nchr,
            "ped": arr(c.dat, PED_DTYPE, c.nid),
            "map": arr(c.genmap, MAP_DTYPE, c.nsnp),
            "chr": arr(c.chrstat, CHR_DTYPE, c.nchr),
            "geno": arr(c.geno, np.dtype("<i4"), npack1 * c.nsnp).reshape(c.nsnp, npack1),
        }
    finally:
        epihost().epi_container_free(ctypes.byref(c))
    return out


def _plink_raise(rc):
    H = epihost()
    raise EpiError(rc, f"{H.epih_strerror(rc).decode()}: {H.epi_plink_error().decode()}")


def _records(ptr, dtype, n):
    buf = (ctypes.c_char * (dtype.itemsize * n)).from_address(ptr)
    return np.frombuffer(bytearray(buf), dtype=dtype)


def read_plink(files, clean=False, impute=False, rnorm=False, seed=1, container=None):
    """PLINK data to the scan's inputs (the reference's data mode, src/readplink.c:1296-1435).

    files: (ped, map) for text PLINK, or a fileset prefix string for prefix.bed/.bim/.fam.
    clean / impute / rnorm: cleangeno, guessmissing and rnorm, in that order, sharing one
    random stream seeded like srand(seed) (seed 1 = the reference's unseeded run).
    container: optional path; the packed container is written there.
    Returns a dict with nid, nsnp, nchr, ped, map, chr records, codes (nsnp x nid uint8,
    classes 0..3), words (layout A1) and the cleaning report."""
    H = epihost()
    d = Plink()
    if isinstance(files, (str, bytes, os.PathLike)):
        pre = os.fsdecode(files)
        rc = H.epi_read_bed(os.fsencode(pre + ".bed"), os.fsencode(pre + ".bim"), os.fsencode(pre + ".fam"),
                            ctypes.byref(d))
    else:
        ped, mp = files
        rc = H.epi_read_ped_map(os.fsencode(ped), os.fsencode(mp), ctypes.byref(d))
    if rc:
        _plink_raise(rc)
    try:
        flags = (PLINK_CLEAN if clean else 0) | (PLINK_IMPUTE if impute else 0) | (PLINK_RNORM if rnorm else 0)
        rep = PlinkReport()
        rc = H.epi_plink_process(ctypes.byref(d), flags, seed, ctypes.byref(rep))
        if rc:
            _plink_raise(rc)
        c = Container()
        rc = H.epi_plink_to_container(ctypes.byref(d), ctypes.byref(c))
        if rc:
            _plink_raise(rc)
        try:
            npack1 = c.npack + (c.remain > 0)
            out = {
                "nid": d.nid, "nsnp": d.nsnp, "nchr": d.nchr,
                "ped": _records(d.dat, PED_DTYPE, d.nid),
                "map": _records(d.genmap, MAP_DTYPE, d.nsnp),
                "chr": _records(d.chrstat, CHR_DTYPE, d.nchr),
                "codes": _records(d.codes, np.dtype(np.uint8), d.nid * d.nsnp).reshape(d.nsnp, d.nid),
                "words": _records(c.geno, np.dtype("<i4"), npack1 * d.nsnp).reshape(d.nsnp, npack1),
                "report": {k: getattr(rep, k) for k, _ in PlinkReport._fields_},
            }
            if container is not None:
                rc = H.epi_write_container(os.fsencode(container), ctypes.byref(c))
                if rc:
                    raise EpiError(rc, H.epih_strerror(rc).decode())
        finally:
            H.epi_container_free(ctypes.byref(c))
    finally:
        H.epi_plink_free(ctypes.byref(d))
    return out


def write_ped_map(container, ped, mp):
    """Container file back to PLINK text (extractfrombinary, src/readplink.c:150-224)."""
    H = epihost()
    c = Container()
    rc = H.epi_read_container(os.fsencode(container), ctypes.byref(c))
    if rc:
        raise EpiError(rc, H.epih_strerror(rc).decode())
    try:
        rc = H.epi_write_ped_map(ctypes.byref(c), os.fsencode(ped), os.fsencode(mp))
        if rc:
            _plink_raise(rc)
    finally:
        H.epi_container_free(ctypes.byref(c))


def make_records(nid, nsnp, chr_sizes, phen):
    """ped/map/chromosome records for a synthetic dataset."""
    ped = np.zeros(nid, dtype=PED_DTYPE)
    for i in range(nid):
        ped[i]["family"] = b"F%d" % i
        ped[i]["id"] = b"I%d" % i
    ped["sex"] = b"1"
    ped["phen"] = phen
    mp = np.zeros(nsnp, dtype=MAP_DTYPE)
    chrs = np.zeros(len(chr_sizes), dtype=CHR_DTYPE)
    start = 0
    for i, sz in enumerate(chr_sizes):
        chrs[i]["chrid"] = i + 1
        chrs[i]["chrname"] = b"%d" % (i + 1)
        chrs[i]["chrsize"] = sz
        chrs[i]["chrstart"] = start
        for s in range(start, start + sz):
            mp[s]["chr"] = b"%d" % (i + 1)
            mp[s]["snpname"] = b"rs%d" % (s + 1)
            mp[s]["pd"] = s - start + 1
        start += sz
    return ped, mp, chrs


def write_container(path, words, nid, ped, mp, chrs):
    words = np.ascontiguousarray(words, dtype=np.int32)
    npack, remain, _ = layout(nid)
    ped = np.ascontiguousarray(ped)
    mp = np.ascontiguousarray(mp)
    chrs = np.ascontiguousarray(chrs)
    c = Container(nid, words.shape[0], npack, remain, len(chrs), ped.ctypes.data, mp.ctypes.data,
                  chrs.ctypes.data, words.ctypes.data)
    rc = epihost().epi_write_container(os.fsencode(path), ctypes.byref(c))
    if rc:
        raise EpiError(rc, epihost().epih_strerror(rc).decode())


def perform_scan(words, nid, ped, mp, chrs, filename, test="f", t1=6.5, t2=10.5, form="mfma", device=0,
                 datafile="data.epigpu", phenfile="", perm=0, quiet=True, safe=False):
    """Run the C host driver (mirror of the reference's perform_scan)."""
    H = epihost()
    o = RunOpts()
    H.epi_run_opts_default(ctypes.byref(o))
    o.test, o.t1, o.t2, o.device, o.form, o.perm = ord(test), t1, t2, device, FORMS[form], perm
    o.safe = int(safe)
    o.datafile, o.phenfile, o.quiet = datafile.encode(), phenfile.encode(), int(quiet)
    rep = RunReport()
    words = np.ascontiguousarray(words, dtype=np.int32)
    npack, remain, _ = layout(nid)
    ped, mp, chrs = (np.ascontiguousarray(x) for x in (ped, mp, chrs))
    rc = H.epi_perform_scan(ctypes.byref(o), _ptr(words), _ptr(chrs), _ptr(ped), _ptr(mp), nid, words.shape[0],
                            npack, remain, len(chrs), os.fsencode(filename), ctypes.byref(rep))
    if rc:
        raise EpiError(rc, H.epih_strerror(rc).decode())
    return {f: getattr(rep, f) for f, _ in RunReport._fields_}


def permute_phen(seed, ped):
    """Phenotypes of permutation `seed` (the reference's permute() with srand(seed))."""
    ped = np.ascontiguousarray(ped)
    out = np.zeros(len(ped), dtype=np.float32)
    epihost().epi_permute_phen(seed, len(ped), _ptr(ped), _ptr(out))
    return out


def permute_order(seed, nid):
    """Record order of permutation `seed`: order[i] = original index of the record at position i."""
    out = np.zeros(nid, dtype=np.int32)
    epihost().epi_permute_order(seed, nid, _ptr(out))
    return out


def write_perm_file(path, ped, order, phen=None):
    """The reference's perm<k>.txt side file for a permutation order (phen by original index, or ped's)."""
    ped = np.ascontiguousarray(ped)
    order = np.ascontiguousarray(order, dtype=np.int32)
    ph = None if phen is None else np.ascontiguousarray(phen, dtype=np.float32)
    rc = epihost().epi_write_perm_file(os.fsencode(path), len(ped), _ptr(ped), _ptr(order), _ptr(ph))
    if rc:
        raise EpiError(rc, epihost().epih_strerror(rc).decode())


def perform_scan_multi(words, nid, mp, chrs, phen, outfiles, perms=None, phenfiles=None, test="f", t1=6.5,
                       t2=10.5, form="mfma", device=0, datafile="data.epigpu", quiet=True, safe=False):
    """Several phenotype vectors (rows of phen) in one batched launch; scan s writes outfiles[s]."""
    H = epihost()
    o = RunOpts()
    H.epi_run_opts_default(ctypes.byref(o))
    o.test, o.t1, o.t2, o.device, o.form = ord(test), t1, t2, device, FORMS[form]
    o.datafile, o.quiet, o.safe = datafile.encode(), int(quiet), int(safe)
    rep = RunReport()
    words = np.ascontiguousarray(words, dtype=np.int32)
    phen = np.ascontiguousarray(np.atleast_2d(phen), dtype=np.float32)
    n = phen.shape[0]
    npack, remain, _ = layout(nid)
    mp, chrs = (np.ascontiguousarray(x) for x in (mp, chrs))
    names = (ctypes.c_char_p * n)(*[os.fsencode(f) for f in outfiles])
    pf = (ctypes.c_char_p * n)(*[(x or "").encode() for x in (phenfiles or [""] * n)])
    pm = np.ascontiguousarray(perms if perms is not None else np.zeros(n), dtype=np.int32)
    rc = H.epi_perform_scan_multi(ctypes.byref(o), _ptr(words), _ptr(chrs), _ptr(mp), nid, words.shape[0], npack,
                                  remain, len(chrs), n, _ptr(phen), _ptr(pm), ctypes.cast(pf, ctypes.c_void_p),
                                  ctypes.cast(names, ctypes.c_void_p), ctypes.byref(rep))
    if rc:
        raise EpiError(rc, H.epih_strerror(rc).decode())
    return {f: getattr(rep, f) for f, _ in RunReport._fields_}


def perform_scan_perms(words, nid, ped, mp, chrs, fmt, first, last, test="f", t1=6.5, t2=10.5, form="mfma",
                       device=0, datafile="data.epigpu", phenfile="", quiet=True):
    """Permutations first..last in one batched launch; output file k = fmt % k."""
    H = epihost()
    o = RunOpts()
    H.epi_run_opts_default(ctypes.byref(o))
    o.test, o.t1, o.t2, o.device, o.form = ord(test), t1, t2, device, FORMS[form]
    o.datafile, o.phenfile, o.quiet = datafile.encode(), phenfile.encode(), int(quiet)
    rep = RunReport()
    words = np.ascontiguousarray(words, dtype=np.int32)
    npack, remain, _ = layout(nid)
    ped, mp, chrs = (np.ascontiguousarray(x) for x in (ped, mp, chrs))
    rc = H.epi_perform_scan_perms(ctypes.byref(o), _ptr(words), _ptr(chrs), _ptr(ped), _ptr(mp), nid,
                                  words.shape[0], npack, remain, len(chrs), first, last, os.fsencode(fmt),
                                  ctypes.byref(rep))
    if rc:
        raise EpiError(rc, H.epih_strerror(rc).decode())
    return {f: getattr(rep, f) for f, _ in RunReport._fields_}


def resume_state(filename, datafile, nid, nsnp, perm=0, test="f", t1=6.5, t2=10.5):
    st = ResumeState()
    rc = epihost().epi_resume(os.fsencode(filename), datafile.encode(), nid, nsnp, perm, ord(test), t1, t2,
                              ctypes.byref(st))
    if rc:
        raise EpiError(rc, epihost().epih_strerror(rc).decode())
    return {f: getattr(st, f) for f, _ in ResumeState._fields_}


def format_row(test, snp1, name1, chr1, snp2, name2, chr2, df1, df2, F, F2=0.0):
    buf = ctypes.create_string_buffer(512)
    epihost().epi_format_row(buf, 512, ord(test), snp1, name1.encode(), chr1.encode(), snp2, name2.encode(),
                             chr2.encode(), df1, df2, F, F2)
    return buf.value.decode()


# ----------------------------------------------------------- device side --

class Scanner:
    """One episcan.h context (one HIP stream on one device)."""

    def __init__(self, device=0):
        self.L = episcan()
        self.ctx = ctypes.c_void_p()
        rc = self.L.epi_open(device, ctypes.byref(self.ctx))
        if rc:
            raise EpiError(rc, "epi_open: " + self.L.epi_strerror(rc).decode())
        self.nid = None
        self._batches = weakref.WeakSet()  # a batch must be destroyed before its context

    def _chk(self, rc, what):
        if rc:
            raise EpiError(rc, f"{what}: {self.L.epi_last_error(self.ctx).decode()}")

    def close(self):
        for b in list(self._batches):
            b.close()
        if self.ctx:
            self.L.epi_close(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        self.L.epi_device_name(self.ctx, buf, 256)
        return buf.value.decode()

    def set_genotypes(self, words, nid):
        words = np.ascontiguousarray(words, dtype=np.int32)
        assert words.shape[1] == layout(nid)[2]
        self._chk(self.L.epi_set_genotypes(self.ctx, _ptr(words), nid, words.shape[0]), "epi_set_genotypes")
        self.nid = nid

    def set_phenotypes(self, phen, nine_is_value=False):
        """nine_is_value: -9 is a phenotype value, not missing (the reference's safe mode)."""
        phen = np.ascontiguousarray(np.atleast_2d(phen), dtype=np.float32)
        self._chk(self.L.epi_set_phenotypes_ex(self.ctx, _ptr(phen), phen.shape[0], phen.shape[1],
                                               PHEN_NINE_IS_VALUE if nine_is_value else 0), "epi_set_phenotypes")

    def set_chromosomes(self, starts, sizes):
        s = np.ascontiguousarray(starts, dtype=np.int32)
        z = np.ascontiguousarray(sizes, dtype=np.int32)
        self._chk(self.L.epi_set_chromosomes(self.ctx, _ptr(s), _ptr(z), len(s)), "epi_set_chromosomes")

    def set_hit_capacity(self, cap):
        self._chk(self.L.epi_set_hit_capacity(self.ctx, cap), "epi_set_hit_capacity")

    def set_max_tiles(self, n):
        """Bound on a single scan's device tile list; longer lists are built and launched in pieces."""
        self._chk(self.L.epi_set_max_tiles(self.ctx, n), "epi_set_max_tiles")

    def schedule(self, form="mfma"):
        sc = Schedule()
        self._chk(self.L.epi_get_schedule(self.ctx, FORMS[form], ctypes.byref(sc)), "epi_get_schedule")
        return sc.as_dict()

    def set_schedule(self, form="mfma", **kw):
        sc = Schedule()
        self._chk(self.L.epi_get_schedule(self.ctx, FORMS[form], ctypes.byref(sc)), "epi_get_schedule")
        for k, v in kw.items():
            setattr(sc, k, v)
        self._chk(self.L.epi_set_schedule(self.ctx, FORMS[form], ctypes.byref(sc)), "epi_set_schedule")

    def autotune(self, test="f", form="mfma", t1=6.5, t2=10.5, phen=0, dims=0, reps=1):
        """Measure candidate single-scan schedules (coordinate descent) and keep the fastest."""
        p = self.params(test, form, t1, t2, phen)
        best, ms, n = Schedule(), ctypes.c_double(), ctypes.c_int32()
        log = (TuneEntry * 64)()
        self._chk(self.L.epi_autotune(self.ctx, ctypes.byref(p), dims, reps, ctypes.byref(best), ctypes.byref(ms),
                                      log, 64, ctypes.byref(n)), "epi_autotune")
        return {"best": best.as_dict(), "best_ms": ms.value,
                "log": [dict(log[i].sched.as_dict(), ms=log[i].ms) for i in range(min(n.value, 64))]}

    @staticmethod
    def params(test="f", form="mfma", t1=6.5, t2=10.5, phen=0, blk_begin=0, blk_end=0, safe=False):
        p = ScanParams()
        p.test, p.form, p.t1, p.t2, p.phen = ord(test), FORMS[form], t1, t2, phen
        p.blk_begin, p.blk_end, p.safe = blk_begin, blk_end, int(safe)
        return p

    def scan(self, test="f", form="mfma", t1=6.5, t2=10.5, phen=0, blk_begin=0, blk_end=0, fetch=True,
             safe=False):
        p = self.params(test, form, t1, t2, phen, blk_begin, blk_end, safe)
        self._chk(self.L.epi_scan(self.ctx, ctypes.byref(p)), "epi_scan")
        n = ctypes.c_int64()
        ovf = ctypes.c_int32()
        self._chk(self.L.epi_hit_count(self.ctx, ctypes.byref(n), ctypes.byref(ovf)), "epi_hit_count")
        ms, pairs, nl = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int32()
        self.L.epi_last_scan_stats(self.ctx, ctypes.byref(ms), ctypes.byref(pairs), ctypes.byref(nl))
        out = {"nhits": n.value, "overflow": bool(ovf.value), "kernel_ms": ms.value, "pairs": pairs.value,
               "launches": nl.value}
        if fetch:
            m = min(n.value, 1 << 40)
            hits = np.zeros(m, dtype=HIT_DTYPE)
            F = np.zeros(m, dtype=np.float64)
            F2 = np.zeros(m, dtype=np.float64)
            blk = np.zeros(m, dtype=np.int32)
            got = ctypes.c_int64()
            self._chk(self.L.epi_get_hits(self.ctx, _ptr(hits), _ptr(F), _ptr(F2), _ptr(blk), m,
                                          ctypes.byref(got)), "epi_get_hits")
            g = got.value
            out.update(hits=hits[:g], F64=F[:g], F2_64=F2[:g], block=blk[:g])
        return out

    def last_tiles(self):
        """Tiles of the last scan, no-allocation scan or dump by kind."""
        f, sp, g = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.L.epi_last_tiles(self.ctx, ctypes.byref(f), ctypes.byref(sp), ctypes.byref(g)),
                  "epi_last_tiles")
        return {"fast": f.value, "sparse": sp.value, "general": g.value}

    def scan_noalloc(self, hits_dev, cap, test="f", form="mfma", t1=6.5, t2=10.5, phen=0, blk_begin=0, blk_end=0,
                     safe=False):
        """epi_scan_noalloc: hits (HIT64_DTYPE records, unsorted) into the device buffer at address hits_dev of
        cap records; returns the total hit count (may exceed cap)."""
        p = self.params(test, form, t1, t2, phen, blk_begin, blk_end, safe)
        n = ctypes.c_int64()
        self._chk(self.L.epi_scan_noalloc(self.ctx, ctypes.byref(p), ctypes.c_void_p(hits_dev), cap,
                                          ctypes.byref(n)), "epi_scan_noalloc")
        return n.value

    def pair_dump(self, a0, a1, b0, b1, test="i", form="mfma", phen=0):
        p = self.params(test, form, 6.5, 10.5, phen)
        na, nb = a1 - a0, b1 - b0
        counts = np.full((na, nb, 9), -1, dtype=np.int32)
        sums = np.full((na, nb, 9), np.nan, dtype=np.float64)
        F = np.full((na, nb), np.nan, dtype=np.float64)
        F2 = np.full((na, nb), np.nan, dtype=np.float64)
        self._chk(self.L.epi_pair_dump(self.ctx, ctypes.byref(p), a0, a1, b0, b1, _ptr(counts), _ptr(sums),
                                       _ptr(F), _ptr(F2)), "epi_pair_dump")
        return counts, sums, F, F2


class Batch:
    """An episcan.h batch: many (dataset, phenotype) scans in one captured persistent launch.

    Holds a reference to its Scanner, whose context must outlive it."""

    def __init__(self, scanner):
        self.sc = scanner
        self.L = scanner.L
        self.b = ctypes.c_void_p()
        scanner._chk(self.L.epi_batch_create(scanner.ctx, ctypes.byref(self.b)), "epi_batch_create")
        scanner._batches.add(self)
        self.nscan = 0

    def _chk(self, rc, what):
        self.sc._chk(rc, what)

    def close(self):
        if self.b:
            self.L.epi_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_dataset(self, words, nid, chr_starts=None, chr_sizes=None):
        words = np.ascontiguousarray(words, dtype=np.int32)
        assert words.shape[1] == layout(nid)[2]
        d = ctypes.c_int32()
        if chr_starts is None:
            s = z = None
            nchr = 0
        else:
            s = np.ascontiguousarray(chr_starts, dtype=np.int32)
            z = np.ascontiguousarray(chr_sizes, dtype=np.int32)
            nchr = len(s)
        self._chk(self.L.epi_batch_add_dataset(self.b, _ptr(words), nid, words.shape[0], _ptr(s), _ptr(z), nchr,
                                               ctypes.byref(d)), "epi_batch_add_dataset")
        return d.value

    def add_phenotypes(self, dataset, phen, nine_is_value=False):
        phen = np.ascontiguousarray(np.atleast_2d(phen), dtype=np.float32)
        first = ctypes.c_int32()
        self._chk(self.L.epi_batch_add_phenotypes_ex(self.b, dataset, _ptr(phen), phen.shape[0],
                                                     PHEN_NINE_IS_VALUE if nine_is_value else 0,
                                                     ctypes.byref(first)), "epi_batch_add_phenotypes")
        self.nscan = first.value + phen.shape[0]
        return list(range(first.value, first.value + phen.shape[0]))

    def set_hit_capacity(self, cap):
        self._chk(self.L.epi_batch_set_hit_capacity(self.b, cap), "epi_batch_set_hit_capacity")

    def scan(self, test="f", form="mfma", t1=6.5, t2=10.5, safe=False):
        p = Scanner.params(test, form, t1, t2, safe=safe)
        self._chk(self.L.epi_batch_scan(self.b, ctypes.byref(p)), "epi_batch_scan")
        return self.stats()

    def stats(self):
        ms, pairs, tiles = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        cap, grid = ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.L.epi_batch_stats(self.b, ctypes.byref(ms), ctypes.byref(pairs), ctypes.byref(tiles),
                                         ctypes.byref(cap), ctypes.byref(grid)), "epi_batch_stats")
        n, ovf = self.hit_count(-1)
        return {"kernel_ms": ms.value, "pairs": pairs.value, "tiles": tiles.value, "captures": cap.value,
                "grid": grid.value, "nhits": n, "overflow": ovf}

    def schedule(self, form="mfma"):
        sc = Schedule()
        self._chk(self.L.epi_batch_get_schedule(self.b, FORMS[form], ctypes.byref(sc)), "epi_batch_get_schedule")
        return sc.as_dict()

    def set_schedule(self, form="mfma", **kw):
        sc = Schedule()
        self._chk(self.L.epi_batch_get_schedule(self.b, FORMS[form], ctypes.byref(sc)), "epi_batch_get_schedule")
        for k, v in kw.items():
            setattr(sc, k, v)
        self._chk(self.L.epi_batch_set_schedule(self.b, FORMS[form], ctypes.byref(sc)), "epi_batch_set_schedule")

    def autotune(self, test="f", form="mfma", t1=6.5, t2=10.5, dims=TUNE_ALL, reps=1):
        """Measure candidate schedules (coordinate descent) and keep the fastest."""
        p = Scanner.params(test, form, t1, t2)
        best, ms, n = Schedule(), ctypes.c_double(), ctypes.c_int32()
        log = (TuneEntry * 64)()
        self._chk(self.L.epi_batch_autotune(self.b, ctypes.byref(p), dims, reps, ctypes.byref(best),
                                            ctypes.byref(ms), log, 64, ctypes.byref(n)), "epi_batch_autotune")
        return {"best": best.as_dict(), "best_ms": ms.value,
                "log": [dict(log[i].sched.as_dict(), ms=log[i].ms) for i in range(min(n.value, 64))]}

    def hit_count(self, scan=-1):
        n, ovf = ctypes.c_int64(), ctypes.c_int32()
        self._chk(self.L.epi_batch_hit_count(self.b, scan, ctypes.byref(n), ctypes.byref(ovf)),
                  "epi_batch_hit_count")
        return n.value, bool(ovf.value)

    def hits(self, scan):
        m, _ = self.hit_count(scan)
        hits = np.zeros(m, dtype=HIT_DTYPE)
        F = np.zeros(m, dtype=np.float64)
        F2 = np.zeros(m, dtype=np.float64)
        blk = np.zeros(m, dtype=np.int32)
        got = ctypes.c_int64()
        self._chk(self.L.epi_batch_get_hits(self.b, scan, _ptr(hits), _ptr(F), _ptr(F2), _ptr(blk), m,
                                            ctypes.byref(got)), "epi_batch_get_hits")
        g = got.value
        return {"nhits": g, "hits": hits[:g], "F64": F[:g], "F2_64": F2[:g], "block": blk[:g]}
