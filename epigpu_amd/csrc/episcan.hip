// libepiscan.so - C ABI over the MI355X (gfx950) pair-scan kernels.
//
// Device state of one dataset (a context holds one; a batch holds many):
//   words   [nsnp][npack1] int32   genotypes as uploaded (layout A1), freed after the transcode
//   mbits   [3][Kw][nsnp]  uint32  bitplanes of classes 0, 1 and call present (MFMA form)
//   bits    [3][Kw][nsnp]  uint32  bitplanes of the two rarest classes and call present (plain form)
//   cls     [nsnp]         int32   plain-form plane -> class map
//   per phenotype r:
//     wts   [8][Kp] int8   V, 3 limbs of q, 4 limbs of round(q^2/2^zbits)
//     qv    [Kp]    int32  fixed-point phenotype q (0 where missing and for outliers)
//     yc    [Kp]    float  centred phenotype scaled by 2^-ybits to |yc| < 1 (0 where missing)
//     vbits [Kw]    uint32 phenotype-present bits
//     marg  [nsnp]  Marg   per-SNP class counts and sums
//     oidx, ocodes, ov     outlier individuals (at most OUT_MAX), their codes per SNP and values
//     mp    [Kp]    MissPhen V, q, z per individual (SPARSE tiles)
//   for SPARSE tiles (built at the first scan that has them, for the chromosome layout it was built for):
//     moff  [nsnp+1] int64, mlist int32   per SNP, the individuals with a missing call
//     vidx  [nsnp]  int32  virtual SNP index: each chromosome starts at a multiple of 64
//     tbits [nw][nid][4] uint64  transposed bitplanes over the virtual order (64 SNPs per word)
// Kp = nid rounded up to 256 (the plain kernel's stage), Kw = Kp / 32; padding
// individuals are absent. The MFMA K loops stop at nid rounded up to 128.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "episcan.h"
#include "kern_batch.h"
#include "kern_mfma.h"
#include "kern_plain.h"
#include "transcode.h"

using namespace epi;

namespace {

constexpr int KPAD = 256;
// The transcode grid has one row of workgroups per 32 individuals, and a grid
// row index stays below 65536: at most 8191 K-pads of individuals.
constexpr int MAX_NID = 65535 / (KPAD / 32) * KPAD;
constexpr long long QMAX = 8355711;     // largest |q| with three balanced int8 limbs
constexpr long long ZMAX = 2139062143;  // largest |z| with four balanced int8 limbs
// The MFMA form's phenotype range. The central SD s_c is the SD of the
// present values without the 2 OUT_MAX farthest from the median (a quarter of
// them when there are fewer than 258). Among the OUT_MAX farthest, values more
// than OUT_SD s_c from the central mean are outliers, carried exactly. The
// encoded values must then lie within BULK_SD s_c of the centre, which bounds
// the fixed-point quantum at BULK_SD / QMAX = 7.7e-6 s_c; a phenotype beyond
// that is refused by MFMA scans (EPI_ERR_RANGE), and the plain form, FP32 per
// value, takes it.
constexpr double OUT_SD = 32.0;
constexpr double BULK_SD = 64.0;

struct PhenDev {
    int8_t *wts = nullptr;
    int32_t *qv = nullptr;
    float *yc = nullptr;
    uint32_t *vbits = nullptr;
    Marg *marg = nullptr;
    int32_t *oidx = nullptr;    // [nout] outlier individuals
    uint32_t *ocodes = nullptr; // [ceil(nout/16)][nsnp], built with the marginals
    OutVals *ov = nullptr;
    MissPhen *mp = nullptr;
    int nout = 0;
    int fbits = 0;  // q = round((y - centre) * 2^fbits)
    int zbits = 0;  // Z limbs hold round(q^2 / 2^zbits)
    int ybits = 0;  // plain form: yc = (y - centre) * 2^-ybits, |yc| < 1
    double centre = 0.0;
    int nV = 0;
    long long qV = 0;
    double sqV = 0.0;  // sum q^2
    double ysV = 0.0;  // sum yc
    double y2V = 0.0;  // sum yc^2
    double ouSum = 0.0, ouSq = 0.0;  // sum of the outliers' values (units 2^-fbits) and of their squares
    double sc = 0.0, maxdev = 0.0;   // central SD; largest |y - centre| of the encoded values
    bool mfma_ok = true;
    std::string mfma_why;
};

struct Dataset {
    int nid = 0, nsnp = 0, npack = 0, remain = 0, npack1 = 0, Kp = 0, Kw = 0;
    int32_t *words = nullptr;
    uint32_t *mbits = nullptr;
    uint32_t *bits = nullptr;
    int32_t *dmiss = nullptr;
    int32_t *cls = nullptr;
    std::vector<int> miss;    // per SNP: missing calls
    std::vector<int> missc;   // prefix count of SNPs with missing calls
    std::vector<int> heavyc;  // prefix count of SNPs with more than sparse_max() missing calls
    std::vector<long long> missum;  // prefix sum of missing calls
    std::vector<PhenDev> ph;
    bool marg_ok = false;
    std::vector<int> chrstart, chrsize, snpchr;
    // SPARSE tiles
    bool sp_ok = false;       // built for the current genotypes and chromosomes
    long long *moff = nullptr;
    int32_t *mlist = nullptr;
    int32_t *vidx = nullptr;
    uint64_t *tbits = nullptr;
    long long tstride = 0;    // uint64 words per 64-SNP column of tbits (4 nid)
};

// Tile lists of a scan by kind: FAST (no missing call), SPARSE (missing
// calls, few per SNP), GENERAL (the rest). Device order: FAST, SPARSE, GENERAL.
struct TileLists {
    std::vector<int4> fast, gen, sp;
    long long pairs = 0;
    size_t count() const { return fast.size() + sp.size() + gen.size(); }
};

}  // namespace

struct epi_ctx {
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    char name[256] = {0};
    Dataset d;
    // hits
    DevHit *hits = nullptr;
    unsigned long long *nhits = nullptr;
    long long cap = 1ll << 22;
    long long last_nhits = 0;
    // tiles; the device list of the last scan is kept and reused while the
    // dataset, block range, form and schedule stay the same
    int4 *tiles = nullptr;
    size_t tiles_cap = 0;
    size_t max_tiles = (size_t)1 << 25;  // 512 MiB of int4
    TileLists tl;
    bool tl_valid = false;
    int tl_bb = 0, tl_be = 0, tl_form = -1;
    epi_schedule tl_sched = {};
    // stats
    double last_ms = 0.0;
    long long last_pairs = 0;
    int last_launches = 0;
    long long last_tiles[3] = {0, 0, 0};  // FAST, SPARSE, GENERAL tiles of the last scan
    epi_schedule sched[2];  // per form
    void *scratch = nullptr;  // this device's no-allocation scratch (address from hipGetSymbolAddress)
};

struct epi_batch {
    epi_ctx *ctx = nullptr;
    std::vector<Dataset *> data;
    std::vector<std::pair<int, int>> scans;  // (dataset, phenotype)
    MfmaScan *d_mscan = nullptr;
    PlainScan *d_pscan = nullptr;
    int4 *d_work = nullptr;
    int *d_qpos = nullptr, *d_qbegin = nullptr, *d_qend = nullptr;  // [NKIND][NQ]: FAST, GENERAL, SPARSE items
    size_t work_cap = 0, scan_cap = 0;
    long long nkind[NKIND] = {0, 0, 0};  // work items of each tile kind (FAST first in d_work)
    DevHit *hits = nullptr;
    unsigned long long *nhits = nullptr;
    long long cap = 1ll << 22;
    // work list of the last build
    epi_schedule sched[2];  // per form
    epi_schedule work_sched = {};
    bool dirty = true;
    int work_form = -1;
    long long nwork = 0, pairs = 0;
    // captured graph of the last scan
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    epi_scan_params gkey = {};
    bool gvalid = false;
    int captures = 0;
    int grid = 0;
    // results
    long long last_nhits = 0;
    double last_ms = 0.0;
    bool fetched = false;
    std::vector<DevHit> hh;
    std::vector<std::vector<int>> by_scan;
};

namespace {

int fail(epi_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(c, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((c), EPI_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                     \
    } while (0)

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
int dalloc(epi_ctx *c, T *&p, size_t count)
{
    dfree(p);
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        return fail(c, EPI_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    return EPI_OK;
}

void free_phen(PhenDev &p)
{
    dfree(p.wts);
    dfree(p.qv);
    dfree(p.yc);
    dfree(p.vbits);
    dfree(p.marg);
    dfree(p.oidx);
    dfree(p.ocodes);
    dfree(p.ov);
    dfree(p.mp);
}

void drop_sparse(Dataset &d)
{
    dfree(d.moff);
    dfree(d.mlist);
    dfree(d.vidx);
    dfree(d.tbits);
    d.sp_ok = false;
}

void drop_phenotypes(Dataset &d)
{
    for (auto &p : d.ph) free_phen(p);
    d.ph.clear();
    d.marg_ok = false;
}

void free_dataset(Dataset &d)
{
    dfree(d.words);
    dfree(d.mbits);
    dfree(d.bits);
    dfree(d.dmiss);
    dfree(d.cls);
    drop_sparse(d);
    for (auto &p : d.ph) free_phen(p);
    d.ph.clear();
}

// ------------------------------------------------------ dataset preparation --

// SNPs with at most this many missing calls can be in SPARSE tiles (the
// per-pair counts of the missing-call sums are 16-bit).
int sparse_max(int nid) { return std::min(65535, std::max(8, nid / 16)); }

// A tile with missing calls is SPARSE (under dense_missing 0) when its SNPs
// miss at most 1 / SPARSE_RATE of their calls on average: measured at
// N = 4000 with 64-row tiles, the per-individual sums of a SPARSE tile cost
// less than the NM rows of a GENERAL tile up to about 0.7 % missing calls
// (0.5 %: 17.7 ms against 20.5 ms for the scan; 1 %: 24.7 against 20.5).
constexpr long long SPARSE_RATE = 150;

int ds_upload_genotypes(epi_ctx *c, Dataset &d, const int32_t *words, int nid, int nsnp)
{
    if (!d.ph.empty() && nid != d.nid) drop_phenotypes(d);  // a new cohort: the old phenotypes no longer apply
    drop_sparse(d);
    d.nid = nid;
    d.nsnp = nsnp;
    d.npack = nid / 16;
    d.remain = nid % 16;
    d.npack1 = d.npack + (d.remain > 0);
    d.Kp = (nid + KPAD - 1) / KPAD * KPAD;
    d.Kw = d.Kp / 32;
    size_t nwords = (size_t)nsnp * d.npack1;
    int rc;
    if ((rc = dalloc(c, d.words, nwords)) || (rc = dalloc(c, d.mbits, (size_t)3 * d.Kw * nsnp)) ||
        (rc = dalloc(c, d.bits, (size_t)3 * d.Kw * nsnp)) || (rc = dalloc(c, d.dmiss, (size_t)nsnp)) ||
        (rc = dalloc(c, d.cls, (size_t)nsnp)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(d.words, words, nwords * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(class_order, dim3((nsnp + 255) / 256), dim3(256), 0, c->stream, (const uint32_t *)d.words, nid,
                       d.npack, d.remain, d.npack1, nsnp, d.cls, d.dmiss);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(transcode_bits, dim3((nsnp + 255) / 256, d.Kw), dim3(256), 0, c->stream,
                       (const uint32_t *)d.words, nid, d.npack, d.remain, d.npack1, nsnp, d.Kw, (const int32_t *)d.cls,
                       d.bits);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(transcode_bits, dim3((nsnp + 255) / 256, d.Kw), dim3(256), 0, c->stream,
                       (const uint32_t *)d.words, nid, d.npack, d.remain, d.npack1, nsnp, d.Kw, (const int32_t *)nullptr,
                       d.mbits);
    HIPCHK(c, hipGetLastError());
    d.miss.assign(nsnp, 0);
    HIPCHK(c, hipMemcpyAsync(d.miss.data(), d.dmiss, (size_t)nsnp * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(d.words);  // the uploaded words only feed the transcode
    d.missc.assign(nsnp + 1, 0);
    d.heavyc.assign(nsnp + 1, 0);
    d.missum.assign(nsnp + 1, 0);
    const int smax = sparse_max(nid);
    for (int s = 0; s < nsnp; s++) {
        d.missc[s + 1] = d.missc[s] + (d.miss[s] != 0);
        d.heavyc[s + 1] = d.heavyc[s] + (d.miss[s] > smax);
        d.missum[s + 1] = d.missum[s] + d.miss[s];
    }
    d.marg_ok = false;
    // chromosome layout resets to one block over all SNPs
    d.chrstart.assign(1, 0);
    d.chrsize.assign(1, nsnp);
    d.snpchr.assign(nsnp, 0);
    return EPI_OK;
}

// New genotypes. A failure after the arguments are checked leaves the
// dataset empty (genotypes and phenotypes unset), never half updated.
int ds_set_genotypes(epi_ctx *c, Dataset &d, const int32_t *words, int nid, int nsnp)
{
    if (!words || nid < 2 || nsnp < 2) return fail(c, EPI_ERR_ARG, "need nid >= 2, nsnp >= 2 and words");
    if (nid > MAX_NID) return fail(c, EPI_ERR_RANGE, "nid %d above the supported %d", nid, MAX_NID);
    d.marg_ok = false;
    const int rc = ds_upload_genotypes(c, d, words, nid, nsnp);
    if (rc) {
        free_dataset(d);
        d = Dataset();
    }
    return rc;
}

// Balanced signed 8-bit limbs of x, least significant first; false if x needs more than n.
bool split_limbs(long long x, int n, int8_t *out)
{
    for (int i = 0; i < n; i++) {
        const long long l = ((x % 256) + 256 + 128) % 256 - 128;
        out[i] = (int8_t)l;
        x = (x - l) / 256;
    }
    return x == 0;
}

// Host-side encoding of one phenotype, before anything is allocated.
struct PhenEnc {
    std::vector<int8_t> w;      // [NWP][Kp]
    std::vector<int32_t> qv;    // [Kp]
    std::vector<float> yc;      // [Kp]
    std::vector<uint32_t> vb;   // [Kw]
    std::vector<int32_t> oidx;  // outlier individuals
    std::vector<MissPhen> mp;   // [Kp]
    OutVals ov = {};
    PhenDev p;                  // scalars only
};

// Phenotype prep (src/EpiOpenCL.c:1364-1376 computes mY and SSY over all nid
// and the kernel drops -9 values, :135-147; here the missing values are
// dropped up front). Centre, outliers and fixed-point scale of the MFMA
// form, the plain form's scaled FP32 values, presence bits.
int encode_phen(epi_ctx *c, const float *y, int nid, int Kp, bool nine_value, int r, PhenEnc &e)
{
    auto missing = [&](float v) { return v != v || (!nine_value && v == -9.0f); };
    std::vector<int> P;
    P.reserve(nid);
    for (int k = 0; k < nid; k++) {
        const float v = y[k];
        if (missing(v)) continue;
        if (std::isinf(v)) return fail(c, EPI_ERR_RANGE, "phenotype %d has an infinite value", r);
        P.push_back(k);
    }
    const int nv = (int)P.size();
    if (nv < 2) return fail(c, EPI_ERR_RANGE, "phenotype %d has fewer than 2 present values", r);
    PhenDev &p = e.p;
    // Outliers among the ntop values farthest from the median (which far values
    // cannot drag), measured against the central values (all but the ncen farthest).
    std::vector<float> med(nv);
    for (int t = 0; t < nv; t++) med[t] = y[P[t]];
    std::nth_element(med.begin(), med.begin() + nv / 2, med.end());
    const double c0 = (double)med[nv / 2];
    const int ncen = std::min(2 * OUT_MAX, (nv - 2) / 4), ntop = std::min(OUT_MAX, ncen);
    std::vector<int> ord(P);
    auto farther = [&](int x, int z) {
        const double dx = std::fabs((double)y[x] - c0), dz = std::fabs((double)y[z] - c0);
        return dx != dz ? dx > dz : x < z;
    };
    if (ncen > 0) {
        std::nth_element(ord.begin(), ord.begin() + ncen - 1, ord.end(), farther);
        std::sort(ord.begin(), ord.begin() + ncen, farther);
    }
    std::vector<char> far(nid, 0), out(nid, 0);
    for (int t = 0; t < ncen; t++) far[ord[t]] = 1;
    double mb = 0.0, sb = 0.0;
    for (int k : P)
        if (!far[k]) mb += (double)y[k];
    mb /= (nv - ncen);
    for (int k : P)
        if (!far[k]) sb += ((double)y[k] - mb) * ((double)y[k] - mb);
    sb = std::sqrt(sb / (nv - ncen));
    for (int t = 0; t < ntop; t++) {
        const int k = ord[t];
        if (std::fabs((double)y[k] - mb) > OUT_SD * sb) {
            out[k] = 1;
            e.oidx.push_back(k);
        }
    }
    p.nout = (int)e.oidx.size();
    double centre = 0.0;
    for (int k : P)
        if (!out[k]) centre += (double)y[k];
    centre /= (nv - p.nout);
    double maxabs = 0.0, maxall = 0.0;
    for (int k : P) {
        const double dv = std::fabs((double)y[k] - centre);
        maxall = std::max(maxall, dv);
        if (!out[k]) maxabs = std::max(maxabs, dv);
    }
    p.sc = sb;
    p.maxdev = maxabs;
    p.mfma_ok = maxabs <= BULK_SD * sb;
    if (!p.mfma_ok) {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "phenotype %d: more than %d values lie far out (after %d outliers, a value %.3g central SDs "
                 "out remains, above %g); the MFMA form's fixed point cannot hold it to 1e-5, use the plain form",
                 r, OUT_MAX, p.nout, sb > 0.0 ? maxabs / sb : INFINITY, BULK_SD);
        p.mfma_why = buf;
    }
    int f = 0;
    if (maxabs > 0.0) f = (int)std::floor(std::log2((double)QMAX / maxabs));
    while (maxabs > 0.0 && std::ldexp(maxabs, f) > (double)QMAX) f--;
    // the plain form's FP32 values are scaled by a power of two to |yc| < 1, so that products of
    // phenotypes far from unit scale neither overflow nor underflow FP32 (F is scale-free)
    int ye = 0;
    if (maxall > 0.0) std::frexp(maxall, &ye);
    const int Kw = Kp / 32;
    e.w.assign((size_t)NWP * Kp, 0);
    e.qv.assign(Kp, 0);
    e.yc.assign(Kp, 0.0f);
    e.vb.assign(Kw, 0u);
    e.mp.assign(Kp, MissPhen{});
    long long qmax2 = 0;
    for (int k : P) {
        const long long q = out[k] ? 0 : std::llround(std::ldexp((double)y[k] - centre, f));
        if (q > QMAX || q < -QMAX) return fail(c, EPI_ERR_RANGE, "fixed-point phenotype out of range");
        qmax2 = std::max(qmax2, q * q);
        e.qv[k] = (int32_t)q;
    }
    // q^2 limbs: the smallest shift that fits four limbs (15 when |q| reaches QMAX)
    int zb = 0;
    while (std::llround(std::ldexp((double)qmax2, -zb)) > ZMAX) zb++;
    long long qsum = 0;
    double sq = 0.0, ys = 0.0, y2 = 0.0;
    for (int k : P) {
        const long long q = e.qv[k];
        const long long z = std::llround(std::ldexp((double)(q * q), -zb));
        int8_t lq[3], lz[4];
        if (!split_limbs(q, 3, lq) || !split_limbs(z, 4, lz)) return fail(c, EPI_ERR_RANGE, "limb overflow");
        e.w[0 * (size_t)Kp + k] = 1;
        for (int t = 0; t < 3; t++) e.w[(1 + t) * (size_t)Kp + k] = lq[t];
        for (int t = 0; t < 4; t++) e.w[(4 + t) * (size_t)Kp + k] = lz[t];
        MissPhen &mk = e.mp[k];
        mk.v = 1;
        mk.q = (int32_t)q;
        mk.z = (int32_t)z;
        mk.qf = (float)q;
        mk.zf = (float)z;
        qsum += q;
        sq += (double)q * (double)q;
        const double dv = (double)y[k] - centre;
        const float ycf = (float)std::ldexp(dv, -ye);
        e.yc[k] = ycf;
        ys += (double)ycf;
        y2 += (double)ycf * (double)ycf;
        e.vb[k >> 5] |= 1u << (k & 31);
    }
    for (int t = 0; t < p.nout; t++) {
        const double u = std::ldexp((double)y[e.oidx[t]] - centre, f);
        e.ov.d[t] = u;
        e.ov.f[t] = (float)u;
        p.ouSum += u;
        p.ouSq += u * u;
    }
    p.fbits = f;
    p.zbits = zb;
    p.ybits = ye;
    p.centre = centre;
    p.nV = nv;
    p.qV = qsum;
    p.sqV = sq;
    p.ysV = ys;
    p.y2V = y2;
    return EPI_OK;
}

// Encode every phenotype on the host, then allocate and upload. On any
// failure the dataset is left without phenotypes (a scan then reports
// EPI_ERR_STATE), never with a partly built set.
int ds_set_phenotypes(epi_ctx *c, Dataset &d, const float *phen, int nphen, int nid, int flags)
{
    const bool nine_value = (flags & EPI_PHEN_NINE_IS_VALUE) != 0;
    if (!phen || nphen < 1 || nid < 2) return fail(c, EPI_ERR_ARG, "need phenotypes");
    if (nid > MAX_NID) return fail(c, EPI_ERR_RANGE, "nid %d above the supported %d", nid, MAX_NID);
    if (d.nid && nid != d.nid) return fail(c, EPI_ERR_ARG, "nid %d differs from genotypes' nid %d", nid, d.nid);
    drop_phenotypes(d);
    const int Kp = (nid + KPAD - 1) / KPAD * KPAD, Kw = Kp / 32;
    std::vector<PhenEnc> enc(nphen);
    for (int r = 0; r < nphen; r++) {
        int rc = encode_phen(c, phen + (size_t)r * nid, nid, Kp, nine_value, r, enc[r]);
        if (rc) return rc;
    }
    std::vector<PhenDev> ph(nphen);
    int rc = EPI_OK;
    for (int r = 0; r < nphen && rc == EPI_OK; r++) {
        PhenEnc &e = enc[r];
        PhenDev &p = ph[r];
        p = e.p;
        if ((rc = dalloc(c, p.wts, e.w.size())) || (rc = dalloc(c, p.qv, (size_t)Kp)) ||
            (rc = dalloc(c, p.yc, (size_t)Kp)) || (rc = dalloc(c, p.vbits, (size_t)Kw)) ||
            (rc = dalloc(c, p.ov, 1)) || (rc = dalloc(c, p.mp, (size_t)Kp)))
            break;
        if (p.nout && (rc = dalloc(c, p.oidx, (size_t)p.nout))) break;
        hipError_t he = hipMemcpyAsync(p.wts, e.w.data(), e.w.size(), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) he = hipMemcpyAsync(p.qv, e.qv.data(), (size_t)Kp * 4, hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) he = hipMemcpyAsync(p.yc, e.yc.data(), (size_t)Kp * 4, hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(p.vbits, e.vb.data(), (size_t)Kw * 4, hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) he = hipMemcpyAsync(p.ov, &e.ov, sizeof(OutVals), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(p.mp, e.mp.data(), (size_t)Kp * sizeof(MissPhen), hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess && p.nout)
            he = hipMemcpyAsync(p.oidx, e.oidx.data(), (size_t)p.nout * 4, hipMemcpyHostToDevice, c->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
        if (he != hipSuccess) rc = fail(c, EPI_ERR_HIP, "phenotype upload failed: %s", hipGetErrorString(he));
    }
    if (rc) {
        for (auto &p : ph) free_phen(p);
        return rc;
    }
    if (!d.nid) {
        d.nid = nid;
        d.Kp = Kp;
        d.Kw = Kw;
    }
    d.ph.swap(ph);
    d.marg_ok = false;
    return EPI_OK;
}

int ds_set_chromosomes(epi_ctx *c, Dataset &d, const int32_t *chrstart, const int32_t *chrsize, int nchr)
{
    if (!chrstart || !chrsize || nchr < 1) return fail(c, EPI_ERR_ARG, "need chromosomes");
    if (!d.nsnp) return fail(c, EPI_ERR_STATE, "set genotypes before chromosomes");
    std::vector<int> snpchr(d.nsnp, -1);
    for (int i = 0; i < nchr; i++) {
        if (chrsize[i] < 0 || chrstart[i] < 0 || (long long)chrstart[i] + chrsize[i] > d.nsnp)
            return fail(c, EPI_ERR_ARG, "chromosome %d out of range", i);
        for (int s = chrstart[i]; s < chrstart[i] + chrsize[i]; s++) {
            if (snpchr[s] >= 0) return fail(c, EPI_ERR_ARG, "chromosomes %d and %d overlap", snpchr[s], i);
            snpchr[s] = i;
        }
    }
    d.chrstart.assign(chrstart, chrstart + nchr);
    d.chrsize.assign(chrsize, chrsize + nchr);
    d.snpchr.swap(snpchr);
    drop_sparse(d);  // its virtual SNP order follows the chromosomes
    return EPI_OK;
}

// Per-SNP marginals and outlier codes of every phenotype (they depend on both
// the genotypes and the phenotypes).
int ds_marginals(epi_ctx *c, Dataset &d)
{
    if (d.marg_ok) return EPI_OK;
    if (!d.mbits) return fail(c, EPI_ERR_STATE, "genotypes not set");
    if (d.ph.empty()) return fail(c, EPI_ERR_STATE, "phenotypes not set");
    for (auto &p : d.ph) {
        int rc = dalloc(c, p.marg, (size_t)d.nsnp);
        if (rc) return rc;
        hipLaunchKernelGGL(marginals_kernel, dim3((d.nsnp + 63) / 64), dim3(256), 0, c->stream,
                           (const uint32_t *)d.mbits, d.Kw, d.nsnp, (const uint32_t *)p.vbits, (const int32_t *)p.qv,
                           (const float *)p.yc, (const int32_t *)d.cls, (const MissPhen *)p.mp, p.marg);
        HIPCHK(c, hipGetLastError());
        if (p.nout) {
            if ((rc = dalloc(c, p.ocodes, (size_t)(p.nout + 15) / 16 * d.nsnp))) return rc;
            hipLaunchKernelGGL(outlier_codes_kernel, dim3((d.nsnp + 255) / 256), dim3(256), 0, c->stream,
                               (const uint32_t *)d.mbits, d.Kw, d.nsnp, (const int32_t *)p.oidx, p.nout, p.ocodes);
            HIPCHK(c, hipGetLastError());
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    d.marg_ok = true;
    return EPI_OK;
}

// Virtual SNP order of the SPARSE tiles: chromosome i starts at a multiple of
// 64, so that the a- or b-range of every tile (which starts at its
// chromosome's start plus a multiple of the tile height) is one 64-SNP word,
// or half of one for 32-row tiles. Returns the number of words.
long long virtual_order(const Dataset &d, std::vector<int32_t> &vidx, std::vector<int32_t> *snp_of_v)
{
    vidx.assign(d.nsnp, 0);
    long long nv = 0;
    for (size_t i = 0; i < d.chrstart.size(); i++) {
        for (int t = 0; t < d.chrsize[i]; t++) vidx[d.chrstart[i] + t] = (int32_t)(nv + t);
        nv += (d.chrsize[i] + 63) / 64 * 64;
    }
    if (nv == 0) nv = 64;
    if (snp_of_v) {
        snp_of_v->assign(nv, -1);
        for (size_t i = 0; i < d.chrstart.size(); i++)
            for (int t = 0; t < d.chrsize[i]; t++) (*snp_of_v)[vidx[d.chrstart[i] + t]] = d.chrstart[i] + t;
    }
    return nv / 64;
}

// Device data of the SPARSE tiles for the current genotypes and chromosomes
// (the miss lists, the virtual SNP order and the transposed bitplanes), built
// when the genotypes or the chromosomes are set, for datasets with missing
// calls.
int ds_sparse(epi_ctx *c, Dataset &d)
{
    drop_sparse(d);
    if (!d.mbits || d.missc.empty() || d.missc[d.nsnp] == 0) return EPI_OK;
    std::vector<int32_t> vidx, sov;
    const long long nw = virtual_order(d, vidx, &sov);
    std::vector<long long> moff(d.nsnp + 1, 0);
    for (int s = 0; s < d.nsnp; s++) moff[s + 1] = moff[s] + d.miss[s];
    d.tstride = 4LL * d.nid;
    int32_t *dsov = nullptr;
    int rc;
    if ((rc = dalloc(c, d.moff, (size_t)d.nsnp + 1)) || (rc = dalloc(c, d.mlist, (size_t)std::max(1LL, moff[d.nsnp]))) ||
        (rc = dalloc(c, d.vidx, (size_t)d.nsnp)) || (rc = dalloc(c, d.tbits, (size_t)nw * d.tstride)) ||
        (rc = dalloc(c, dsov, sov.size()))) {
        dfree(dsov);
        drop_sparse(d);
        return rc;
    }
    hipError_t he = hipMemcpyAsync(d.moff, moff.data(), moff.size() * sizeof(long long), hipMemcpyHostToDevice,
                                   c->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(d.vidx, vidx.data(), vidx.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(dsov, sov.data(), sov.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(miss_list_kernel, dim3((d.nsnp + 255) / 256), dim3(256), 0, c->stream,
                           (const uint32_t *)d.mbits, d.Kw, d.nsnp, d.nid, (const long long *)d.moff, d.mlist);
        he = hipGetLastError();
    }
    if (he == hipSuccess) {
        hipLaunchKernelGGL(tbits_kernel, dim3((unsigned)nw, (d.nid + 31) / 32), dim3(64), 0, c->stream,
                           (const uint32_t *)d.mbits, d.Kw, d.nsnp, d.nid, (const int32_t *)dsov, (int)nw, d.tbits);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    dfree(dsov);
    if (he != hipSuccess) {
        drop_sparse(d);
        return fail(c, EPI_ERR_HIP, "SPARSE tile data: %s", hipGetErrorString(he));
    }
    d.sp_ok = true;
    return EPI_OK;
}

// SPARSE tiles serve MFMA scans of datasets whose SPARSE data are built; the
// schedule can turn them off.
bool use_sparse(const Dataset &d, int form, const epi_schedule &sc)
{
    return form == EPI_FORM_MFMA && sc.dense_missing != 1 && d.sp_ok;
}

// MFMA scans of phenotype r need a phenotype the fixed point holds to 1e-5.
int check_mfma_phen(epi_ctx *c, const Dataset &d, int r)
{
    if (r >= 0 && r < (int)d.ph.size() && !d.ph[r].mfma_ok) return fail(c, EPI_ERR_RANGE, "%s", d.ph[r].mfma_why.c_str());
    return EPI_OK;
}

MfmaScan mfma_scan(const Dataset &d, int r, int tag)
{
    const PhenDev &p = d.ph[r];
    MfmaScan S;
    S.bits = d.mbits;
    S.wts = p.wts;
    S.marg = p.marg;
    S.nsnp = d.nsnp;
    S.Kp = d.Kp;
    S.Kl = (d.nid + KC - 1) / KC * KC;
    S.nV = p.nV;
    S.tag = tag;
    S.qV = p.qV;
    S.qVd = (double)p.qV;
    S.sqV = p.sqV;
    S.zunit = std::ldexp(1.0, p.zbits);
    S.ocodes = p.ocodes;
    S.ov = p.ov;
    S.nout = p.nout;
    S.ouSum = p.ouSum;
    S.ouSq = p.ouSq;
    S.tbits = d.tbits;
    S.vidx = d.vidx;
    S.moff = d.moff;
    S.mlist = d.mlist;
    S.mp = p.mp;
    S.tstride = d.tstride;
    return S;
}

PlainScan plain_scan(const Dataset &d, int r, int tag)
{
    const PhenDev &p = d.ph[r];
    PlainScan S;
    S.bits = d.bits;
    S.yc = p.yc;
    S.vbits = p.vbits;
    S.marg = p.marg;
    S.nsnp = d.nsnp;
    S.Kw = d.Kw;
    S.nV = p.nV;
    S.tag = tag;
    S.ysV = p.ysV;
    S.y2V = p.y2V;
    return S;
}

// ------------------------------------------------------- no allocation --
// Scratch of epi_scan_noalloc: the hit counter in the first 256 bytes, then
// the tile list. Each device holds its own copy until the process ends. Host
// code reaches it only through the address hipGetSymbolAddress returns on
// that device (scratch_of), never through the array's name, which on the
// host is a stand-in and no device address.
constexpr size_t NOALLOC_BYTES = (size_t)16 << 20;
constexpr size_t NOALLOC_TILES = (NOALLOC_BYTES - 256) / sizeof(int4);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex scratch_mu;
std::vector<void *> scratch_ptr;  // per device ordinal

// This device's scratch address, looked up once per device (the device must be current).
int scratch_of(epi_ctx *c, void **out)
{
    std::lock_guard<std::mutex> g(scratch_mu);
    if ((int)scratch_ptr.size() <= c->device) scratch_ptr.resize(c->device + 1, nullptr);
    if (!scratch_ptr[c->device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, EPI_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        scratch_ptr[c->device] = ptr;
    }
    *out = scratch_ptr[c->device];
    return EPI_OK;
}

// Clears the hit counter with a kernel (no memset node in a captured stream).
__global__ void zero_counter(unsigned long long *n) { *n = 0ull; }

// Schedules. FAST tiles (no missing call in either SNP range) are
// fast_rows x 64. GENERAL tiles are fast_rows x 32 in the MFMA form (the
// same workgroup size as its FAST tiles, two waves per SIMD) and 16 x 64 in
// the plain form.
constexpr int TILE_B = 64;

int gen_rows(int form, const epi_schedule &s) { return form == EPI_FORM_MFMA ? s.fast_rows : 16; }
int gen_cols(int form, const epi_schedule &) { return form == EPI_FORM_MFMA ? 32 : 64; }

epi_schedule default_schedule(int form)
{
    epi_schedule s;
    memset(&s, 0, sizeof(s));
    // 32-row MFMA tiles keep two workgroups per CU (measured faster than 64
    // rows at N=4000 with missing calls and at N=8000 without)
    s.fast_rows = 32;
    s.super_a = 8;
    s.super_b = 4;
    s.queues = 8;
    s.order = 0;
    return s;
}

bool valid_schedule(int form, const epi_schedule &s)
{
    bool rows = form == EPI_FORM_MFMA ? (s.fast_rows == 64 || s.fast_rows == 32)
                                      : (s.fast_rows == 32 || s.fast_rows == 16);
    return rows && s.super_a >= 1 && s.super_a <= 64 && s.super_b >= 1 && s.super_b <= 64 &&
           (s.queues == 1 || s.queues == 8) && s.order >= 0 && s.order <= 2 &&
           (s.dense_missing >= 0 && s.dense_missing <= 2);
}

bool same_schedule(const epi_schedule &x, const epi_schedule &y)
{
    return x.fast_rows == y.fast_rows && x.super_a == y.super_a && x.super_b == y.super_b && x.queues == y.queues &&
           x.order == y.order && x.dense_missing == y.dense_missing;
}

void emit_tile(const Dataset &d, int a0, int a1, int b0, int b1, int grows, int gcols, bool sparse, bool sparse_all,
               TileLists &out)
{
    auto anymiss = [&](int s0, int s1) { return d.missc[s1] - d.missc[s0] > 0; };
    auto anyheavy = [&](int s0, int s1) { return d.heavyc[s1] - d.heavyc[s0] > 0; };
    if (a1 - 1 <= b0) return;  // no pair with a > b
    if (!anymiss(a0, a1) && !anymiss(b0, b1)) {
        out.fast.push_back(make_int4(a0, a1, b0, b1));
        return;
    }
    const long long nmiss = d.missum[a1] - d.missum[a0] + d.missum[b1] - d.missum[b0];
    if (sparse && !anyheavy(a0, a1) && !anyheavy(b0, b1) &&
        (sparse_all || nmiss * SPARSE_RATE <= (long long)(a1 - a0 + b1 - b0) * d.nid)) {
        out.sp.push_back(make_int4(a0, a1, b0, b1));
        return;
    }
    for (int x = a0; x < a1; x += grows) {
        int x1 = std::min(a1, x + grows);
        for (int y = b0; y < b1; y += gcols) {
            int y1 = std::min(b1, y + gcols);
            if (x1 - 1 <= y) continue;
            out.gen.push_back(make_int4(x, x1, y, y1));
        }
    }
}

// Where a piecewise tile build resumes: the next chr x chr block (index
// i(i+1)/2 + j) and, within it, the next super-tile row.
struct TileCursor {
    int blk = 0;
    int sa = 0;
};

// Tiles over the chr x chr blocks [cur.blk, blk_end). Within a block the
// tiles are emitted in super-tiles of super_a x super_b tiles so that
// neighbouring workgroups share SNP data in L2. The build stops after the
// super-tile row in which the list reaches max_tiles (a soft bound: one row
// of super-tiles is never split) and leaves cur at the next row; it returns
// true once the block range is exhausted. A block's pairs are counted with
// its first row.
bool build_tiles(const Dataset &d, int blk_end, int form, const epi_schedule &sc, size_t max_tiles, TileCursor &cur,
                 TileLists &out)
{
    const int TA = sc.fast_rows, TB = TILE_B, grows = gen_rows(form, sc), gcols = gen_cols(form, sc);
    const int nchr = (int)d.chrstart.size();
    const int SA = sc.super_a, SB = sc.super_b;
    const bool sparse = use_sparse(d, form, sc), sparse_all = sc.dense_missing == 2;
    for (int i = 0; i < nchr; i++) {
        for (int j = 0; j <= i; j++) {
            int blk = i * (i + 1) / 2 + j;
            if (blk < cur.blk) continue;
            if (blk >= blk_end) return true;
            cur.blk = blk;
            int A0 = d.chrstart[i], A1 = A0 + d.chrsize[i];
            int B0 = d.chrstart[j], B1 = B0 + d.chrsize[j];
            if (A1 <= A0 || B1 <= B0) {
                cur.sa = 0;
                continue;
            }
            if (cur.sa == 0)
                for (long long a = A0; a < A1; a++) {  // pairs in this block (a > b)
                    long long lo = B0, hi = std::min<long long>(B1, a);
                    if (hi > lo) out.pairs += hi - lo;
                }
            int nta = (A1 - A0 + TA - 1) / TA, ntb = (B1 - B0 + TB - 1) / TB;
            for (int sa = cur.sa; sa < nta; sa += SA) {
                for (int sb = 0; sb < ntb; sb += SB)
                    for (int ta = sa; ta < std::min(nta, sa + SA); ta++)
                        for (int tb = sb; tb < std::min(ntb, sb + SB); tb++) {
                            int a0 = A0 + ta * TA, a1 = std::min(A1, a0 + TA);
                            int b0 = B0 + tb * TB, b1 = std::min(B1, b0 + TB);
                            emit_tile(d, a0, a1, b0, b1, grows, gcols, sparse, sparse_all, out);
                        }
                if (out.count() >= max_tiles && sa + SA < nta) {
                    cur.sa = sa + SA;
                    return false;
                }
            }
            cur.sa = 0;
            if (out.count() >= max_tiles) {
                cur.blk = blk + 1;
                return blk + 1 >= blk_end || blk + 1 >= nchr * (nchr + 1) / 2;
            }
        }
    }
    return true;
}

// The lists of one scan into dst, in device order (FAST, SPARSE, GENERAL).
int copy_tiles(epi_ctx *c, int4 *dst, const TileLists &tl)
{
    size_t o = 0;
    for (const std::vector<int4> *v : {&tl.fast, &tl.sp, &tl.gen}) {
        if (!v->empty())
            HIPCHK(c, hipMemcpyAsync(dst + o, v->data(), v->size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
        o += v->size();
    }
    return EPI_OK;
}

int upload_tiles(epi_ctx *c, const TileLists &tl)
{
    c->tl_valid = false;
    size_t need = tl.count();
    if (need > c->tiles_cap) {
        int rc = dalloc(c, c->tiles, need);
        if (rc) return rc;
        c->tiles_cap = need;
    }
    return copy_tiles(c, c->tiles, tl);
}

HitSink make_sink(DevHit *hits, unsigned long long *nhits, long long cap, const epi_scan_params *p)
{
    HitSink hs;
    hs.hits = hits;
    hs.nhits = nhits;
    hs.cap = cap;
    hs.test = p->test;
    hs.t1 = p->t1;
    hs.t2 = p->t2;
    hs.t1s = p->t1 * 0.98f;
    hs.t2s = p->t2 * 0.98f;
    hs.tag = p->phen;
    hs.safe = p->safe ? 1 : 0;
    return hs;
}

int check_common(epi_ctx *c, const epi_scan_params *p)
{
    if (!p) return fail(c, EPI_ERR_ARG, "null params");
    if (p->test != 'f' && p->test != 'i') return fail(c, EPI_ERR_ARG, "test must be 'f' or 'i'");
    if (p->form != EPI_FORM_PLAIN && p->form != EPI_FORM_MFMA) return fail(c, EPI_ERR_ARG, "unknown form");
    if (!(p->t1 == p->t1) || !(p->t2 == p->t2)) return fail(c, EPI_ERR_ARG, "threshold is NaN");
    return EPI_OK;
}

int check_params(epi_ctx *c, const epi_scan_params *p)
{
    int rc = check_common(c, p);
    if (rc) return rc;
    if (!c->d.mbits) return fail(c, EPI_ERR_STATE, "genotypes not set");
    if (c->d.ph.empty()) return fail(c, EPI_ERR_STATE, "phenotypes not set");
    if (p->phen < 0 || p->phen >= (int)c->d.ph.size()) return fail(c, EPI_ERR_ARG, "phenotype index out of range");
    if (c->d.chrstart.empty()) return fail(c, EPI_ERR_STATE, "chromosomes not set");
    if (p->form == EPI_FORM_MFMA) return check_mfma_phen(c, c->d, p->phen);
    return EPI_OK;
}

// A launch covers at most 2^31 work-items (the runtime counts the global size in 32 bits; beyond 2^32 it
// would launch only the wrapped remainder), so a long tile list goes out as several launches, each over its
// own slice of the list and a whole number of 256-workgroup groups (the XCD mapping of xcd_tile).
#define LAUNCH_TILES(KERNEL, THREADS, ARGS, FIRST, N)                                                   \
    do {                                                                                                 \
        const int chunk_ = (int)((1u << 31) / (unsigned)(THREADS)) & ~255;                              \
        for (int t0_ = 0; t0_ < (N); t0_ += chunk_) {                                                    \
            (ARGS).tiles = (FIRST) + t0_;                                                                \
            (ARGS).ntiles = std::min(chunk_, (N) - t0_);                                                 \
            hipLaunchKernelGGL(KERNEL, dim3((ARGS).ntiles), dim3(THREADS), 0, c->stream, (ARGS));        \
            HIPCHK(c, hipGetLastError());                                                                \
            c->last_launches++;                                                                          \
        }                                                                                                \
    } while (0)

// Device buffers of a scan's pair kernels: the tile list, the hit records and the hit counter.
struct ScanBufs {
    int4 *tiles;
    DevHit *hits;
    unsigned long long *nhits;
    long long cap;
};

ScanBufs ctx_bufs(const epi_ctx *c) { return {c->tiles, c->hits, c->nhits, c->cap}; }

// Launch the pair kernels of one scan over the uploaded tile lists (fast tiles first).
template <bool DUMP>
int launch_pairs(epi_ctx *c, const epi_scan_params *p, const epi_schedule &sc, const TileLists &tl,
                 const DumpSink &ds, const ScanBufs &B)
{
    const Dataset &d = c->d;
    if (tl.count() > (size_t)INT32_MAX)
        return fail(c, EPI_ERR_RANGE, "a super-tile row of this scan exceeds 2^31 tiles");
    const int nf = (int)tl.fast.size(), ns = (int)tl.sp.size(), ng = (int)tl.gen.size();
    c->last_launches = 0;
    c->last_tiles[0] += nf;
    c->last_tiles[1] += ns;
    c->last_tiles[2] += ng;
    const HitSink hs = make_sink(B.hits, B.nhits, B.cap, p);
    if (p->form == EPI_FORM_MFMA) {
        MfmaArgs A;
        A.S = mfma_scan(d, p->phen, p->phen);
        A.hs = hs;
        A.ds = ds;
        if (ns && !d.sp_ok) return fail(c, EPI_ERR_STATE, "SPARSE tile data not built");
        if (sc.fast_rows == 64) {
            LAUNCH_TILES((mfma_fast_kernel<64, DUMP>), fast_threads<64>(), A, B.tiles, nf);
            LAUNCH_TILES((mfma_sparse_kernel<64, DUMP>), sparse_threads<64>(), A, B.tiles + nf, ns);
            LAUNCH_TILES((mfma_general_kernel<64, DUMP>), general_threads<64>(), A, B.tiles + nf + ns, ng);
        } else {
            LAUNCH_TILES((mfma_fast_kernel<32, DUMP>), fast_threads<32>(), A, B.tiles, nf);
            LAUNCH_TILES((mfma_sparse_kernel<32, DUMP>), sparse_threads<32>(), A, B.tiles + nf, ns);
            LAUNCH_TILES((mfma_general_kernel<32, DUMP>), general_threads<32>(), A, B.tiles + nf + ns, ng);
        }
    } else {
        PlainArgs P;
        P.S = plain_scan(d, p->phen, p->phen);
        P.hs = hs;
        P.ds = ds;
        if (sc.fast_rows == 32)
            LAUNCH_TILES((plain_kernel<2, 4, true, DUMP>), 256, P, B.tiles, nf);
        else
            LAUNCH_TILES((plain_kernel<1, 4, true, DUMP>), 256, P, B.tiles, nf);
        LAUNCH_TILES((plain_kernel<1, 4, false, DUMP>), 256, P, B.tiles + nf + ns, ng);
    }
    return EPI_OK;
}

// Hits of one scan sorted by (block, snp1, snp2) and written out.
void emit_hits(const Dataset &d, const std::vector<DevHit> &h, const std::vector<int> &idx, epi_hit *hits, double *F64,
               double *F2_64, int32_t *block, long long cap, long long *n)
{
    auto blk_of = [&](const DevHit &x) {
        int i = d.snpchr[x.snp1], j = d.snpchr[x.snp2];
        return i * (i + 1) / 2 + j;
    };
    std::vector<int> key(idx);
    std::sort(key.begin(), key.end(), [&](int x, int y) {
        const DevHit &p = h[x], &q = h[y];
        int bp = blk_of(p), bq = blk_of(q);
        if (bp != bq) return bp < bq;
        if (p.snp1 != q.snp1) return p.snp1 < q.snp1;
        if (p.snp2 != q.snp2) return p.snp2 < q.snp2;
        return p.tag < q.tag;
    });
    long long m = std::min<long long>((long long)key.size(), cap);
    for (long long t = 0; t < m; t++) {
        const DevHit &x = h[key[t]];
        if (hits) {
            hits[t].snp1 = x.snp1;
            hits[t].snp2 = x.snp2;
            hits[t].df1 = x.df1;
            hits[t].df2 = x.df2;
            hits[t].F = (float)x.F;
            hits[t].F2 = (float)x.F2;
        }
        if (F64) F64[t] = x.F;
        if (F2_64) F2_64[t] = x.F2;
        if (block) block[t] = blk_of(x);
    }
    if (n) *n = m;
}

// ------------------------------------------------------------------ batch --

// Work items of every scan and the queue boundaries of the schedule.
// Per dataset the items are in locality order: tile-major (super-tiles),
// phenotype-minor, one list per tile kind. order 1 runs the GENERAL and
// SPARSE lists before the FAST one, order 2 takes datasets by decreasing Kp,
// so that the longest items do not end up in the tail. With 8 queues the list
// is cut into NQ contiguous pieces of equal estimated cost, one per XCD;
// with 1 queue every workgroup pulls from one shared queue.
int build_batch_work(epi_batch *b, int form, const epi_schedule &sched)
{
    epi_ctx *c = b->ctx;
    const int nd = (int)b->data.size();
    std::vector<std::vector<int4>> wk[NKIND];
    std::vector<std::vector<double>> ct[NKIND];
    for (int kind = 0; kind < NKIND; kind++) {
        wk[kind].resize(nd);
        ct[kind].resize(nd);
    }
    b->pairs = 0;
    // estimated cost per pair and individual, by kind
    const double unit[NKIND] = {form == EPI_FORM_MFMA ? 16.0 : 4.0, form == EPI_FORM_MFMA ? 40.0 : 10.0, 22.0};
    for (int di = 0; di < nd; di++) {
        const Dataset &d = *b->data[di];
        std::vector<int> sc;
        for (int s = 0; s < (int)b->scans.size(); s++)
            if (b->scans[s].first == di) sc.push_back(s);
        if (sc.empty()) continue;
        TileLists tl;
        TileCursor all;
        build_tiles(d, 1 << 30, form, sched, SIZE_MAX, all, tl);
        b->pairs += tl.pairs * (long long)sc.size();
        for (int kind = 0; kind < NKIND; kind++) {
            const std::vector<int4> &v = kind == 0 ? tl.fast : kind == 1 ? tl.gen : tl.sp;
            for (const int4 &t : v) {
                int ext = (t.y - t.x) | ((t.w - t.z) << 8) | (kind << 16);
                double cst = (double)(t.y - t.x) * (t.w - t.z) * d.Kp * unit[kind] + 5e4;
                for (int s : sc) {
                    wk[kind][di].push_back(make_int4(t.x, t.z, ext, s));
                    ct[kind][di].push_back(cst);
                }
            }
        }
    }
    std::vector<int> dorder(nd);
    for (int i = 0; i < nd; i++) dorder[i] = i;
    if (sched.order == 2)
        std::stable_sort(dorder.begin(), dorder.end(),
                         [&](int x, int y) { return b->data[x]->Kp > b->data[y]->Kp; });
    // One list per tile kind (FAST items first in the work array), each run by its own persistent
    // kernel so that a kernel holds the registers of one tile kind only; within a list the datasets
    // follow the schedule's order. Each list is cut into its own queues.
    std::vector<int4> work;
    std::vector<int> qbegin(NKIND * NQ), qend(NKIND * NQ);
    for (int kind = 0; kind < NKIND; kind++) {
        const int base = (int)work.size();
        std::vector<double> cost;
        for (int di : dorder) {
            work.insert(work.end(), wk[kind][di].begin(), wk[kind][di].end());
            cost.insert(cost.end(), ct[kind][di].begin(), ct[kind][di].end());
        }
        if (work.size() > (size_t)INT32_MAX / 2) return fail(c, EPI_ERR_RANGE, "batch has too many tiles");
        const int n = (int)work.size();
        b->nkind[kind] = n - base;
        int *qb = qbegin.data() + kind * NQ, *qe = qend.data() + kind * NQ;
        for (int q = 0; q < NQ; q++) qb[q] = qe[q] = n;
        qb[0] = base;
        if (sched.queues == 1) {
            qe[0] = n;
        } else {
            double total = 0.0;
            for (double x : cost) total += x;
            double acc = 0.0;
            int q = 0;
            for (int i = base; i < n && q < NQ - 1; i++) {
                acc += cost[i - base];
                if (acc >= total * (q + 1) / NQ) {
                    qe[q] = i + 1;
                    qb[++q] = i + 1;
                }
            }
        }
    }
    b->nwork = (long long)work.size();
    int rc;
    if (work.size() > b->work_cap) {
        if ((rc = dalloc(c, b->d_work, work.size()))) return rc;
        b->work_cap = work.size();
    }
    if (!work.empty())
        HIPCHK(c, hipMemcpyAsync(b->d_work, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice,
                                 c->stream));
    HIPCHK(c, hipMemcpyAsync(b->d_qbegin, qbegin.data(), NKIND * NQ * sizeof(int), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(b->d_qend, qend.data(), NKIND * NQ * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // scan descriptors
    size_t ns = b->scans.size();
    if (ns > b->scan_cap) {
        if ((rc = dalloc(c, b->d_mscan, ns)) || (rc = dalloc(c, b->d_pscan, ns))) return rc;
        b->scan_cap = ns;
    }
    std::vector<MfmaScan> ms(ns);
    std::vector<PlainScan> ps(ns);
    for (size_t s = 0; s < ns; s++) {
        const Dataset &d = *b->data[b->scans[s].first];
        ms[s] = mfma_scan(d, b->scans[s].second, (int)s);
        ps[s] = plain_scan(d, b->scans[s].second, (int)s);
    }
    HIPCHK(c, hipMemcpyAsync(b->d_mscan, ms.data(), ns * sizeof(MfmaScan), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b->d_pscan, ps.data(), ns * sizeof(PlainScan), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b->work_form = form;
    b->work_sched = sched;
    b->dirty = false;
    return EPI_OK;
}

bool same_key(const epi_scan_params &x, const epi_scan_params &y)
{
    return x.test == y.test && x.form == y.form && x.t1 == y.t1 && x.t2 == y.t2 && (x.safe != 0) == (y.safe != 0);
}

void drop_graph(epi_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->gvalid = false;
}

// Threads per workgroup of batch_kernel(form, sc, kind).
int batch_threads_of(int form, const epi_schedule &sc, int kind)
{
    if (form != EPI_FORM_MFMA) return 256;
    if (sc.fast_rows == 64)
        return kind == 2 ? batch_threads<64, 2>() : kind ? batch_threads<64, 1>() : batch_threads<64, 0>();
    return kind == 2 ? batch_threads<32, 2>() : kind ? batch_threads<32, 1>() : batch_threads<32, 0>();
}

const void *batch_kernel(int form, const epi_schedule &sc, int kind)
{
    if (form == EPI_FORM_MFMA) {
        if (sc.fast_rows == 64)
            return kind == 2 ? (const void *)mfma_batch_kernel<64, 2>
                   : kind    ? (const void *)mfma_batch_kernel<64, 1>
                             : (const void *)mfma_batch_kernel<64, 0>;
        return kind == 2 ? (const void *)mfma_batch_kernel<32, 2>
               : kind    ? (const void *)mfma_batch_kernel<32, 1>
                         : (const void *)mfma_batch_kernel<32, 0>;
    }
    // the plain form has no SPARSE tiles
    if (sc.fast_rows == 32)
        return kind ? (const void *)plain_batch_kernel<2, 1> : (const void *)plain_batch_kernel<2, 0>;
    return kind ? (const void *)plain_batch_kernel<1, 1> : (const void *)plain_batch_kernel<1, 0>;
}

// Capture queue reset + the persistent scan kernels as one graph: one kernel
// per tile kind that has items (GENERAL and SPARSE first under order 1), each with a grid
// of its occupancy times the CU count, so that every workgroup is resident and
// pulls work until its queues run dry.
int capture_batch(epi_batch *b, const epi_scan_params *p)
{
    epi_ctx *c = b->ctx;
    drop_graph(b);
    BatchArgs B[NKIND];
    const void *kfn[NKIND];
    int grid[NKIND] = {0, 0, 0}, nt[NKIND];
    for (int kind = 0; kind < NKIND; kind++) {
        nt[kind] = batch_threads_of(p->form, b->work_sched, kind);
        B[kind].mscan = b->d_mscan;
        B[kind].pscan = b->d_pscan;
        B[kind].work = b->d_work;
        B[kind].qpos = b->d_qpos + kind * NQ;
        B[kind].qend = b->d_qend + kind * NQ;
        B[kind].hs = make_sink(b->hits, b->nhits, b->cap, p);
        kfn[kind] = batch_kernel(p->form, b->work_sched, kind);
        int per_cu = 1;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn[kind], nt[kind], 0));
        grid[kind] = std::max(1, per_cu) * c->ncu;
    }
    static const int korder[2][NKIND] = {{0, 2, 1}, {1, 2, 0}};
    const int *ko = korder[b->work_sched.order == 1 ? 1 : 0];
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, c->stream, b->d_qpos, (const int *)b->d_qbegin, b->nhits);
    hipError_t le = hipGetLastError();
    for (int t = 0; t < NKIND && le == hipSuccess; t++) {
        const int kind = ko[t];
        if (b->nkind[kind] == 0) continue;
        void *args[] = {&B[kind]};
        le = hipLaunchKernel(kfn[kind], dim3(grid[kind]), dim3(nt[kind]), args, 0, c->stream);
    }
    hipError_t ce = hipStreamEndCapture(c->stream, &b->graph);
    if (le != hipSuccess) return fail(c, EPI_ERR_HIP, "launch during capture: %s", hipGetErrorString(le));
    if (ce != hipSuccess) return fail(c, EPI_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    HIPCHK(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->gkey = *p;
    b->gvalid = true;
    b->grid = 0;
    for (int kind = 0; kind < NKIND; kind++)
        if (b->nkind[kind]) b->grid = std::max(b->grid, grid[kind]);
    b->captures++;
    return EPI_OK;
}

// Work list and graph for (p, the batch's schedule of p->form), rebuilt only
// when the batch, the form, the schedule or the scan settings changed.
int batch_prepare(epi_batch *b, const epi_scan_params *p)
{
    int rc;
    if (p->form == EPI_FORM_MFMA)
        for (const auto &x : b->scans)
            if ((rc = check_mfma_phen(b->ctx, *b->data[x.first], x.second))) return rc;
    const epi_schedule &sc = b->sched[p->form];
    if (b->dirty || b->work_form != p->form || !same_schedule(b->work_sched, sc)) {
        drop_graph(b);
        if ((rc = build_batch_work(b, p->form, sc))) return rc;
    }
    if (!b->gvalid || !same_key(b->gkey, *p))
        if ((rc = capture_batch(b, p))) return rc;
    return EPI_OK;
}

// One replay of the captured graph; device time in ms.
int batch_run(epi_batch *b, double *ms_out)
{
    epi_ctx *c = b->ctx;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long nh = 0;
    HIPCHK(c, hipMemcpyAsync(&nh, b->nhits, sizeof(nh), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    b->last_ms = ms;
    b->last_nhits = (long long)nh;
    b->fetched = false;
    if (ms_out) *ms_out = ms;
    return EPI_OK;
}

int batch_fetch(epi_batch *b)
{
    if (b->fetched) return EPI_OK;
    epi_ctx *c = b->ctx;
    long long have = std::min(b->last_nhits, b->cap);
    b->hh.resize((size_t)have);
    if (have) {
        HIPCHK(c, hipMemcpyAsync(b->hh.data(), b->hits, (size_t)have * sizeof(DevHit), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    b->by_scan.assign(b->scans.size(), {});
    for (long long t = 0; t < have; t++) {
        int s = b->hh[t].tag;
        if (s >= 0 && s < (int)b->scans.size()) b->by_scan[s].push_back((int)t);
    }
    b->fetched = true;
    return EPI_OK;
}

// One scan of the block range of p with schedule sc: tile list, upload,
// pair kernels, hit count. Records the device time and counts in c. A tile
// list of up to c->max_tiles tiles is built once and kept on the device; a
// longer one is built, uploaded and launched in pieces of about that many
// tiles, and the timed region then also holds the uploads of the later pieces.
int run_scan(epi_ctx *c, const epi_scan_params *p, const epi_schedule &sc)
{
    int rc;
    int nchr = (int)c->d.chrstart.size();
    int nblk = nchr * (nchr + 1) / 2;
    int bb = std::max(0, p->blk_begin), be = p->blk_end <= 0 ? nblk : std::min(nblk, p->blk_end);
    TileLists &tl = c->tl;
    const bool cached =
        c->tl_valid && c->tl_bb == bb && c->tl_be == be && c->tl_form == p->form && same_schedule(c->tl_sched, sc);
    TileCursor cur;
    cur.blk = bb;
    bool done = cached;
    if (!cached) {
        tl = TileLists();
        done = build_tiles(c->d, be, p->form, sc, c->max_tiles, cur, tl);
        if ((rc = upload_tiles(c, tl))) return rc;
        if (done) {
            c->tl_valid = true;
            c->tl_bb = bb;
            c->tl_be = be;
            c->tl_form = p->form;
            c->tl_sched = sc;
        }
    }
    HIPCHK(c, hipMemsetAsync(c->nhits, 0, sizeof(unsigned long long), c->stream));
    DumpSink ds;
    memset(&ds, 0, sizeof(ds));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    long long pairs = tl.pairs;
    c->last_tiles[0] = c->last_tiles[1] = c->last_tiles[2] = 0;
    int launches = 0;
    for (;;) {
        if ((rc = launch_pairs<false>(c, p, sc, tl, ds, ctx_bufs(c)))) return rc;
        launches += c->last_launches;
        if (done) break;
        // the launched kernels read the device tile list that the next piece overwrites
        HIPCHK(c, hipStreamSynchronize(c->stream));
        tl = TileLists();
        done = build_tiles(c->d, be, p->form, sc, c->max_tiles, cur, tl);
        if ((rc = upload_tiles(c, tl))) return rc;
        pairs += tl.pairs;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long nh = 0;
    HIPCHK(c, hipMemcpyAsync(&nh, c->nhits, sizeof(nh), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_pairs = pairs;
    c->last_launches = launches;
    c->last_nhits = (long long)nh;
    return EPI_OK;
}

// Coordinate descent over the schedule dimensions in `dims`: each dimension
// in turn is set to every candidate value (others fixed at the best so far)
// and the fastest is kept. measure(s, ms) times one schedule; repeated
// schedules are timed once. Candidates that measure equal keep the earlier.
template <class Measure>
int tune_descent(int form, int dims, epi_schedule &cur, double &tcur, int &nmeasured, epi_tune_entry *log,
                 int log_cap, Measure measure)
{
    std::vector<std::pair<epi_schedule, double>> seen;
    auto timed = [&](const epi_schedule &s, double &ms) -> int {
        for (auto &e : seen)
            if (same_schedule(e.first, s)) {
                ms = e.second;
                return EPI_OK;
            }
        int r = measure(s, ms);
        if (r) return r;
        seen.push_back({s, ms});
        if (log && nmeasured < log_cap) {
            log[nmeasured].sched = s;
            log[nmeasured].ms = ms;
        }
        nmeasured++;
        return EPI_OK;
    };
    int rc = timed(cur, tcur);
    if (rc) return rc;
    auto tryit = [&](const epi_schedule &s) -> int {
        double t = 0.0;
        int r = timed(s, t);
        if (r) return r;
        if (t < tcur) {
            cur = s;
            tcur = t;
        }
        return EPI_OK;
    };
    const int rows_m[] = {32, 64}, rows_p[] = {32, 16};
    const int sup[][2] = {{8, 4}, {4, 4}, {16, 2}, {4, 8}, {1, 1}};
    if ((dims & EPI_TUNE_MISSING) && form == EPI_FORM_MFMA)
        for (int v : {0, 1, 2}) {
            epi_schedule s = cur;
            s.dense_missing = v;
            if ((rc = tryit(s))) return rc;
        }
    if (dims & EPI_TUNE_ROWS)
        for (int v : (form == EPI_FORM_MFMA ? rows_m : rows_p)) {
            epi_schedule s = cur;
            s.fast_rows = v;
            if ((rc = tryit(s))) return rc;
        }
    if (dims & EPI_TUNE_SUPER)
        for (auto &v : sup) {
            epi_schedule s = cur;
            s.super_a = v[0];
            s.super_b = v[1];
            if ((rc = tryit(s))) return rc;
        }
    if (dims & EPI_TUNE_QUEUES)
        for (int v : {8, 1}) {
            epi_schedule s = cur;
            s.queues = v;
            if ((rc = tryit(s))) return rc;
        }
    if (dims & EPI_TUNE_ORDER)
        for (int v : {0, 1, 2}) {
            epi_schedule s = cur;
            s.order = v;
            if ((rc = tryit(s))) return rc;
        }
    return EPI_OK;
}

}  // namespace

// ============================================================== C ABI ====

extern "C" {

const char *epi_strerror(int code)
{
    switch (code) {
    case EPI_OK: return "ok";
    case EPI_ERR_ARG: return "invalid argument";
    case EPI_ERR_HIP: return "HIP runtime error";
    case EPI_ERR_NODEV: return "no such device";
    case EPI_ERR_STATE: return "required data not set";
    case EPI_ERR_NOMEM: return "device out of memory";
    case EPI_ERR_RANGE: return "value out of range";
    default: return "unknown error";
    }
}

const char *epi_last_error(const epi_ctx *c) { return c ? c->err.c_str() : "null context"; }

int epi_open(int device, epi_ctx **out)
{
    if (!out) return EPI_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return EPI_ERR_NODEV;
    if (device < 0 || device >= n) return EPI_ERR_NODEV;
    epi_ctx *c = new epi_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return EPI_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        snprintf(c->name, sizeof(c->name), "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
        c->ncu = prop.multiProcessorCount;
    }
    if (c->ncu <= 0) c->ncu = 1;
    for (int f = 0; f < 2; f++) c->sched[f] = default_schedule(f);
    if (dalloc(c, c->nhits, 1) || dalloc(c, c->hits, (size_t)c->cap)) {
        epi_close(c);
        return EPI_ERR_NOMEM;
    }
    if (scratch_of(c, &c->scratch)) {
        epi_close(c);
        return EPI_ERR_HIP;
    }
    *out = c;
    return EPI_OK;
}

void epi_close(epi_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_dataset(c->d);
    dfree(c->hits);
    dfree(c->nhits);
    dfree(c->tiles);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int epi_device_name(const epi_ctx *c, char *buf, size_t n)
{
    if (!c || !buf || n == 0) return EPI_ERR_ARG;
    snprintf(buf, n, "%s", c->name);
    return EPI_OK;
}

// Marginals as soon as genotypes and phenotypes are both set, so that scans
// (epi_scan_noalloc among them) find them ready. A failure drops the
// phenotypes.
int ready_marginals(epi_ctx *c)
{
    if (!c->d.mbits || c->d.ph.empty()) return EPI_OK;
    const int rc = ds_marginals(c, c->d);
    if (rc) drop_phenotypes(c->d);
    return rc;
}

int epi_set_genotypes(epi_ctx *c, const int32_t *words, int32_t nid, int32_t nsnp)
{
    if (!c) return EPI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    c->last_nhits = 0;
    c->tl_valid = false;
    int rc = ds_set_genotypes(c, c->d, words, nid, nsnp);
    if (!rc) rc = ds_sparse(c, c->d);
    return rc ? rc : ready_marginals(c);
}

int epi_set_phenotypes_ex(epi_ctx *c, const float *phen, int32_t nphen, int32_t nid, int32_t flags)
{
    if (!c) return EPI_ERR_ARG;
    if (flags & ~EPI_PHEN_NINE_IS_VALUE) return fail(c, EPI_ERR_ARG, "unknown phenotype flags");
    HIPCHK(c, hipSetDevice(c->device));
    const int rc = ds_set_phenotypes(c, c->d, phen, nphen, nid, flags);
    return rc ? rc : ready_marginals(c);
}

int epi_set_phenotypes(epi_ctx *c, const float *phen, int32_t nphen, int32_t nid)
{
    return epi_set_phenotypes_ex(c, phen, nphen, nid, 0);
}

int epi_phen_encoding(const float *phen, int32_t nid, int32_t flags, epi_phen_info *info)
{
    if (!phen || !info || nid < 2 || nid > MAX_NID || (flags & ~EPI_PHEN_NINE_IS_VALUE)) return EPI_ERR_ARG;
    PhenEnc e;
    const int Kp = (nid + KPAD - 1) / KPAD * KPAD;
    const int rc = encode_phen(nullptr, phen, nid, Kp, (flags & EPI_PHEN_NINE_IS_VALUE) != 0, 0, e);
    if (rc) return rc;
    memset(info, 0, sizeof(*info));
    info->present = e.p.nV;
    info->outliers = e.p.nout;
    info->mfma_ok = e.p.mfma_ok ? 1 : 0;
    info->fbits = e.p.fbits;
    info->zbits = e.p.zbits;
    info->centre = e.p.centre;
    info->central_sd = e.p.sc;
    info->max_dev = e.p.maxdev;
    for (int t = 0; t < e.p.nout; t++) info->outlier_index[t] = e.oidx[t];
    return EPI_OK;
}

int epi_set_chromosomes(epi_ctx *c, const int32_t *chrstart, const int32_t *chrsize, int32_t nchr)
{
    if (!c) return EPI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    c->tl_valid = false;
    int rc = ds_set_chromosomes(c, c->d, chrstart, chrsize, nchr);
    if (!rc) rc = ds_sparse(c, c->d);
    return rc;
}

void epi_scan_params_default(epi_scan_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->test = 'f';
    p->form = EPI_FORM_MFMA;
    p->t1 = 6.5f;
    p->t2 = 10.5f;
    p->phen = 0;
    p->blk_begin = 0;
    p->blk_end = 0;
}

int epi_set_hit_capacity(epi_ctx *c, int64_t cap)
{
    if (!c || cap < 1) return EPI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = dalloc(c, c->hits, (size_t)cap);
    if (rc) return rc;
    c->cap = cap;
    c->last_nhits = 0;
    return EPI_OK;
}

int epi_set_max_tiles(epi_ctx *c, int64_t max_tiles)
{
    if (!c || max_tiles < 1 || max_tiles > (int64_t)INT32_MAX / 2) return EPI_ERR_ARG;
    c->max_tiles = (size_t)max_tiles;
    c->tl_valid = false;
    return EPI_OK;
}

int epi_scan(epi_ctx *c, const epi_scan_params *p)
{
    if (!c) return EPI_ERR_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ds_marginals(c, c->d))) return rc;
    return run_scan(c, p, c->sched[p->form]);
}

int epi_scan_noalloc(epi_ctx *c, const epi_scan_params *p, epi_hit64 *hits, int64_t cap, int64_t *nhits)
{
    if (!c) return EPI_ERR_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!hits || cap < 0 || !nhits) return fail(c, EPI_ERR_ARG, "need a device hit buffer, its capacity and nhits");
    if (!c->d.marg_ok) return fail(c, EPI_ERR_STATE, "genotypes and phenotypes must both be set");
    if (!c->scratch) return fail(c, EPI_ERR_STATE, "no scratch address for this device");
    HIPCHK(c, hipSetDevice(c->device));
    const epi_schedule &sc = c->sched[p->form];
    const int nchr = (int)c->d.chrstart.size(), nblk = nchr * (nchr + 1) / 2;
    const int bb = std::max(0, p->blk_begin), be = p->blk_end <= 0 ? nblk : std::min(nblk, p->blk_end);
    TileLists tl;
    TileCursor cur;
    cur.blk = bb;
    const bool done = build_tiles(c->d, be, p->form, sc, NOALLOC_TILES + 1, cur, tl);
    const size_t nt = tl.count();
    if (!done || nt > NOALLOC_TILES)
        return fail(c, EPI_ERR_RANGE,
                    "the scan needs more than the %zu tiles the no-allocation scratch holds (%zu MiB); use epi_scan",
                    NOALLOC_TILES, NOALLOC_BYTES >> 20);
    unsigned char *base = static_cast<unsigned char *>(c->scratch);
    const ScanBufs B = {reinterpret_cast<int4 *>(base + 256), reinterpret_cast<DevHit *>(hits),
                        reinterpret_cast<unsigned long long *>(base), (long long)cap};
    if ((rc = copy_tiles(c, B.tiles, tl))) return rc;
    c->last_tiles[0] = c->last_tiles[1] = c->last_tiles[2] = 0;
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, c->stream, B.nhits);
    HIPCHK(c, hipGetLastError());
    DumpSink ds;
    memset(&ds, 0, sizeof(ds));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if ((rc = launch_pairs<false>(c, p, sc, tl, ds, B))) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long nh = 0;
    HIPCHK(c, hipMemcpyAsync(&nh, B.nhits, sizeof(nh), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_pairs = tl.pairs;
    *nhits = (int64_t)nh;
    return EPI_OK;
}

int epi_autotune(epi_ctx *c, const epi_scan_params *p, int32_t dims, int32_t reps, epi_schedule *best,
                 double *best_ms, epi_tune_entry *log, int32_t log_cap, int32_t *nlog)
{
    if (!c) return EPI_ERR_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (reps < 1 || reps > 100) return fail(c, EPI_ERR_ARG, "reps must be in [1, 100]");
    // queues and order do not apply to single scans; the missing-call tiles only with missing calls
    const bool anymiss = !c->d.missc.empty() && c->d.missc[c->d.nsnp] > 0;
    if (dims == 0) dims = EPI_TUNE_ROWS | EPI_TUNE_SUPER | (anymiss ? EPI_TUNE_MISSING : 0);
    dims &= EPI_TUNE_ROWS | EPI_TUNE_SUPER | EPI_TUNE_MISSING;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ds_marginals(c, c->d))) return rc;
    epi_schedule cur = c->sched[p->form];
    double tcur = 0.0;
    int n = 0;
    rc = tune_descent(p->form, dims, cur, tcur, n, log, log_cap, [&](const epi_schedule &s, double &ms) -> int {
        int r = run_scan(c, p, s);  // warm-up
        if (r) return r;
        std::vector<double> t(reps);
        for (int i = 0; i < reps; i++) {
            if ((r = run_scan(c, p, s))) return r;
            t[i] = c->last_ms;
        }
        std::sort(t.begin(), t.end());
        ms = t[reps / 2];
        return EPI_OK;
    });
    if (rc) return rc;
    c->sched[p->form] = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = tcur;
    if (nlog) *nlog = n;
    return EPI_OK;
}

int epi_hit_count(epi_ctx *c, int64_t *nhits, int32_t *overflow)
{
    if (!c) return EPI_ERR_ARG;
    if (nhits) *nhits = c->last_nhits;
    if (overflow) *overflow = c->last_nhits > c->cap;
    return EPI_OK;
}

int epi_get_hits(epi_ctx *c, epi_hit *hits, double *F64, double *F2_64, int32_t *block, int64_t cap, int64_t *n)
{
    if (!c || cap < 0) return EPI_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    long long have = std::min<long long>(c->last_nhits, c->cap);
    std::vector<DevHit> h((size_t)have);
    if (have) {
        HIPCHK(c, hipMemcpyAsync(h.data(), c->hits, (size_t)have * sizeof(DevHit), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    std::vector<int> idx((size_t)have);
    for (long long t = 0; t < have; t++) idx[t] = (int)t;
    long long m = 0;
    emit_hits(c->d, h, idx, hits, F64, F2_64, block, cap, &m);
    if (n) *n = m;
    return EPI_OK;
}

int epi_set_schedule(epi_ctx *c, int32_t form, const epi_schedule *s)
{
    if (!c || !s || (form != EPI_FORM_PLAIN && form != EPI_FORM_MFMA)) return EPI_ERR_ARG;
    if (!valid_schedule(form, *s)) return fail(c, EPI_ERR_ARG, "invalid schedule for this form");
    c->sched[form] = *s;
    return EPI_OK;
}

int epi_get_schedule(epi_ctx *c, int32_t form, epi_schedule *s)
{
    if (!c || !s || (form != EPI_FORM_PLAIN && form != EPI_FORM_MFMA)) return EPI_ERR_ARG;
    *s = c->sched[form];
    return EPI_OK;
}

int epi_last_scan_stats(epi_ctx *c, double *kernel_ms, int64_t *pairs, int32_t *launches)
{
    if (!c) return EPI_ERR_ARG;
    if (kernel_ms) *kernel_ms = c->last_ms;
    if (pairs) *pairs = c->last_pairs;
    if (launches) *launches = c->last_launches;
    return EPI_OK;
}

int epi_last_tiles(epi_ctx *c, int64_t *fast, int64_t *sparse, int64_t *general)
{
    if (!c) return EPI_ERR_ARG;
    if (fast) *fast = c->last_tiles[0];
    if (sparse) *sparse = c->last_tiles[1];
    if (general) *general = c->last_tiles[2];
    return EPI_OK;
}

int epi_pair_dump(epi_ctx *c, const epi_scan_params *p, int32_t a0, int32_t a1, int32_t b0, int32_t b1,
                  int32_t *counts, double *sums, double *F, double *F2)
{
    if (!c) return EPI_ERR_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (a0 < 0 || b0 < 0 || a1 > c->d.nsnp || b1 > c->d.nsnp || a1 <= a0 || b1 <= b0)
        return fail(c, EPI_ERR_ARG, "dump rectangle out of range");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ds_marginals(c, c->d))) return rc;
    const epi_schedule &sc = c->sched[p->form];
    const int TA = sc.fast_rows, TB = TILE_B;
    // Tiles over the rectangle in global SNP space (no block restriction). A
    // tile can be SPARSE where its a- and b-ranges each lie in one chromosome
    // and start on the virtual word grid, as every tile of a scan does.
    const Dataset &d = c->d;
    const bool sp = use_sparse(d, p->form, sc);
    std::vector<int32_t> vidx;
    if (sp) virtual_order(d, vidx, nullptr);
    auto grid = [&](int s0, int s1, int w) {
        const int ch = d.snpchr[s0];
        return ch >= 0 && d.snpchr[s1 - 1] == ch && vidx[s0] % w == 0;
    };
    TileLists all;
    for (int x = a0; x < a1; x += TA)
        for (int y = b0; y < b1; y += TB) {
            const int x1 = std::min(a1, x + TA), y1 = std::min(b1, y + TB);
            emit_tile(d, x, x1, y, y1, gen_rows(p->form, sc), gen_cols(p->form, sc),
                      sp && grid(x, x1, TA) && grid(y, y1, 64), sc.dense_missing == 2, all);
        }
    if ((rc = upload_tiles(c, all))) return rc;
    long long np = (long long)(a1 - a0) * (b1 - b0);
    int32_t *dc = nullptr;
    double *dsum = nullptr, *dF = nullptr, *dF2 = nullptr;
    if ((rc = dalloc(c, dc, (size_t)np * 9)) || (rc = dalloc(c, dsum, (size_t)np * 9)) ||
        (rc = dalloc(c, dF, (size_t)np)) || (rc = dalloc(c, dF2, (size_t)np))) {
        dfree(dc);
        dfree(dsum);
        dfree(dF);
        return rc;
    }
    DumpSink ds;
    ds.counts = dc;
    ds.sums = dsum;
    ds.F = dF;
    ds.F2 = dF2;
    ds.a0 = a0;
    ds.a1 = a1;
    ds.b0 = b0;
    ds.b1 = b1;
    const PhenDev &ph = c->d.ph[p->phen];
    ds.unit = p->form == EPI_FORM_MFMA ? std::ldexp(1.0, -ph.fbits) : std::ldexp(1.0, ph.ybits);
    ds.centre = ph.centre;
    // Pre-fill with the caller's buffers so that pairs with a <= b keep their values.
    int hrc = EPI_OK;
    auto h2d = [&](void *d, const void *hsrc, size_t bytes) {
        if (hrc == EPI_OK && hsrc && hipMemcpyAsync(d, hsrc, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            hrc = fail(c, EPI_ERR_HIP, "dump prefill copy failed");
    };
    h2d(dc, counts, (size_t)np * 9 * 4);
    h2d(dsum, sums, (size_t)np * 9 * 8);
    h2d(dF, F, (size_t)np * 8);
    h2d(dF2, F2, (size_t)np * 8);
    c->last_tiles[0] = c->last_tiles[1] = c->last_tiles[2] = 0;
    if (hrc == EPI_OK) hrc = launch_pairs<true>(c, p, sc, all, ds, ctx_bufs(c));
    auto d2h = [&](void *hdst, const void *d, size_t bytes) {
        if (hrc == EPI_OK && hdst && hipMemcpyAsync(hdst, d, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            hrc = fail(c, EPI_ERR_HIP, "dump copy back failed");
    };
    d2h(counts, dc, (size_t)np * 9 * 4);
    d2h(sums, dsum, (size_t)np * 9 * 8);
    d2h(F, dF, (size_t)np * 8);
    d2h(F2, dF2, (size_t)np * 8);
    if (hrc == EPI_OK && hipStreamSynchronize(c->stream) != hipSuccess) hrc = fail(c, EPI_ERR_HIP, "dump sync failed");
    dfree(dc);
    dfree(dsum);
    dfree(dF);
    dfree(dF2);
    return hrc;
}

// ------------------------------------------------------------------ batch --

int epi_batch_create(epi_ctx *c, epi_batch **out)
{
    if (!c || !out) return EPI_ERR_ARG;
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    epi_batch *b = new epi_batch();
    b->ctx = c;
    for (int f = 0; f < 2; f++) b->sched[f] = default_schedule(f);
    int rc;
    if ((rc = dalloc(c, b->d_qpos, NKIND * NQ)) || (rc = dalloc(c, b->d_qbegin, NKIND * NQ)) ||
        (rc = dalloc(c, b->d_qend, NKIND * NQ)) ||
        (rc = dalloc(c, b->nhits, 1)) || (rc = dalloc(c, b->hits, (size_t)b->cap))) {
        epi_batch_destroy(b);
        return rc;
    }
    *out = b;
    return EPI_OK;
}

void epi_batch_destroy(epi_batch *b)
{
    if (!b) return;
    epi_ctx *c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graph(b);
    for (Dataset *d : b->data) {
        free_dataset(*d);
        delete d;
    }
    dfree(b->d_mscan);
    dfree(b->d_pscan);
    dfree(b->d_work);
    dfree(b->d_qpos);
    dfree(b->d_qbegin);
    dfree(b->d_qend);
    dfree(b->hits);
    dfree(b->nhits);
    delete b;
}

int epi_batch_add_dataset(epi_batch *b, const int32_t *words, int32_t nid, int32_t nsnp, const int32_t *chrstart,
                          const int32_t *chrsize, int32_t nchr, int32_t *dataset)
{
    if (!b) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    Dataset *d = new Dataset();
    int rc = ds_set_genotypes(c, *d, words, nid, nsnp);
    if (!rc && chrstart) rc = ds_set_chromosomes(c, *d, chrstart, chrsize, nchr);
    if (!rc) rc = ds_sparse(c, *d);
    if (rc) {
        free_dataset(*d);
        delete d;
        return rc;
    }
    b->data.push_back(d);
    b->dirty = true;
    if (dataset) *dataset = (int32_t)b->data.size() - 1;
    return EPI_OK;
}

int epi_batch_add_phenotypes(epi_batch *b, int32_t dataset, const float *phen, int32_t nphen, int32_t *first_scan)
{
    return epi_batch_add_phenotypes_ex(b, dataset, phen, nphen, 0, first_scan);
}

int epi_batch_add_phenotypes_ex(epi_batch *b, int32_t dataset, const float *phen, int32_t nphen, int32_t flags,
                                int32_t *first_scan)
{
    if (!b) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    if (flags & ~EPI_PHEN_NINE_IS_VALUE) return fail(c, EPI_ERR_ARG, "unknown phenotype flags");
    if (dataset < 0 || dataset >= (int)b->data.size()) return fail(c, EPI_ERR_ARG, "no such dataset");
    Dataset &d = *b->data[dataset];
    if (!d.ph.empty()) return fail(c, EPI_ERR_STATE, "dataset %d already has phenotypes", dataset);
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ds_set_phenotypes(c, d, phen, nphen, d.nid, flags);
    if (!rc) rc = ds_marginals(c, d);
    if (rc) {
        drop_phenotypes(d);
        return rc;
    }
    if (first_scan) *first_scan = (int32_t)b->scans.size();
    for (int r = 0; r < nphen; r++) b->scans.push_back({dataset, r});
    b->dirty = true;
    return EPI_OK;
}

int epi_batch_set_hit_capacity(epi_batch *b, int64_t cap)
{
    if (!b || cap < 1) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = dalloc(c, b->hits, (size_t)cap);
    if (rc) return rc;
    b->cap = cap;
    b->last_nhits = 0;
    b->fetched = false;
    drop_graph(b);
    return EPI_OK;
}

int epi_batch_scan(epi_batch *b, const epi_scan_params *p)
{
    if (!b) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    int rc = check_common(c, p);
    if (rc) return rc;
    if (b->scans.empty()) return fail(c, EPI_ERR_STATE, "batch has no scans");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = batch_prepare(b, p))) return rc;
    return batch_run(b, nullptr);
}

int epi_batch_set_schedule(epi_batch *b, int32_t form, const epi_schedule *s)
{
    if (!b || !s || (form != EPI_FORM_PLAIN && form != EPI_FORM_MFMA)) return EPI_ERR_ARG;
    if (!valid_schedule(form, *s)) return fail(b->ctx, EPI_ERR_ARG, "invalid schedule for this form");
    b->sched[form] = *s;
    return EPI_OK;
}

int epi_batch_get_schedule(epi_batch *b, int32_t form, epi_schedule *s)
{
    if (!b || !s || (form != EPI_FORM_PLAIN && form != EPI_FORM_MFMA)) return EPI_ERR_ARG;
    *s = b->sched[form];
    return EPI_OK;
}

int epi_batch_autotune(epi_batch *b, const epi_scan_params *p, int32_t dims, int32_t reps, epi_schedule *best,
                       double *best_ms, epi_tune_entry *log, int32_t log_cap, int32_t *nlog)
{
    if (!b) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    int rc = check_common(c, p);
    if (rc) return rc;
    if (b->scans.empty()) return fail(c, EPI_ERR_STATE, "batch has no scans");
    if (reps < 1 || reps > 100) return fail(c, EPI_ERR_ARG, "reps must be in [1, 100]");
    if (dims == 0) dims = EPI_TUNE_ALL;
    HIPCHK(c, hipSetDevice(c->device));
    const int form = p->form;
    epi_schedule cur = b->sched[form];
    double tcur = 0.0;
    int n = 0;
    rc = tune_descent(form, dims, cur, tcur, n, log, log_cap, [&](const epi_schedule &s, double &ms) -> int {
        b->sched[form] = s;
        int r = batch_prepare(b, p);
        if (r) return r;
        if ((r = batch_run(b, nullptr))) return r;  // warm-up
        std::vector<double> t(reps);
        for (int i = 0; i < reps; i++)
            if ((r = batch_run(b, &t[i]))) return r;
        std::sort(t.begin(), t.end());
        ms = t[reps / 2];
        return EPI_OK;
    });
    if (rc) return rc;
    b->sched[form] = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = tcur;
    if (nlog) *nlog = n;
    return EPI_OK;
}

int epi_batch_hit_count(epi_batch *b, int32_t scan, int64_t *nhits, int32_t *overflow)
{
    if (!b) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    if (scan < -1 || scan >= (int)b->scans.size()) return fail(c, EPI_ERR_ARG, "no such scan");
    if (overflow) *overflow = b->last_nhits > b->cap;
    if (scan < 0) {
        if (nhits) *nhits = b->last_nhits;
        return EPI_OK;
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc = batch_fetch(b);
    if (rc) return rc;
    if (nhits) *nhits = (int64_t)b->by_scan[scan].size();
    return EPI_OK;
}

int epi_batch_get_hits(epi_batch *b, int32_t scan, epi_hit *hits, double *F64, double *F2_64, int32_t *block,
                       int64_t cap, int64_t *n)
{
    if (!b || cap < 0) return EPI_ERR_ARG;
    epi_ctx *c = b->ctx;
    if (scan < 0 || scan >= (int)b->scans.size()) return fail(c, EPI_ERR_ARG, "no such scan");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = batch_fetch(b);
    if (rc) return rc;
    long long m = 0;
    emit_hits(*b->data[b->scans[scan].first], b->hh, b->by_scan[scan], hits, F64, F2_64, block, cap, &m);
    if (n) *n = m;
    return EPI_OK;
}

int epi_batch_stats(epi_batch *b, double *kernel_ms, int64_t *pairs, int64_t *tiles, int32_t *captures,
                    int32_t *grid)
{
    if (!b) return EPI_ERR_ARG;
    if (grid) *grid = b->grid;
    if (kernel_ms) *kernel_ms = b->last_ms;
    if (pairs) *pairs = b->pairs;
    if (tiles) *tiles = b->nwork;
    if (captures) *captures = b->captures;
    return EPI_OK;
}

}  // extern "C"
