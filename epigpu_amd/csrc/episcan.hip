// libepiscan.so - C ABI over the MI355X (gfx950) pair-scan kernels.
//
// Device state of one context:
//   words   [nsnp][npack1] int32   genotypes as uploaded (layout A1)
//   planes  [nsnp][3][Kp]  uint8   class masks 0xFF/0x00 for classes 0, 1, NM
//   bits    [3][Kw][nsnp]  uint32  class bitplanes (plain form)
//   per phenotype r:
//     wts   [7][Kp] int8   V, 3 limbs of q, 3 limbs of round(q^2/2^23)
//     qv    [Kp]    int32  fixed-point phenotype q (0 where missing)
//     yc    [Kp]    float  centred phenotype (0 where missing)
//     vbits [Kw]    uint32 phenotype-present bits
//     marg  [nsnp]  Marg   per-SNP class counts and sums
// Kp = nid rounded up to 256, Kw = Kp / 32; padding individuals are absent.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/episcan.h"
#include "kern_mfma.h"
#include "kern_plain.h"

using namespace epi;

namespace {

constexpr int KPAD = 256;
constexpr long long QMAX = 8355711;  // largest |q| with three balanced int8 limbs
constexpr double ZSHIFT = 8388608.0; // 2^23

struct PhenDev {
    int8_t *wts = nullptr;
    int32_t *qv = nullptr;
    float *yc = nullptr;
    uint32_t *vbits = nullptr;
    Marg *marg = nullptr;
    int fbits = 0;        // q = round((y - centre) * 2^fbits)
    double centre = 0.0;
    int nV = 0;
    long long qV = 0;
    double sqV = 0.0;     // sum q^2
    double ysV = 0.0;     // sum yc
    double y2V = 0.0;     // sum yc^2
};

}  // namespace

struct epi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    char name[256] = {0};
    // genotypes
    int nid = 0, nsnp = 0, npack = 0, remain = 0, npack1 = 0, Kp = 0, Kw = 0;
    int32_t *words = nullptr;
    uint8_t *planes = nullptr;
    uint32_t *bits = nullptr;
    int32_t *dmiss = nullptr;
    int32_t *cls = nullptr;  // per SNP plane -> class map of the plain form
    std::vector<int> miss;   // per SNP: any missing call
    std::vector<int> missc;  // prefix count of SNPs with missing calls
    // phenotypes
    int nphen = 0;
    std::vector<PhenDev> ph;
    bool marg_ok = false;
    // chromosomes
    std::vector<int> chrstart, chrsize;
    std::vector<int> snpchr;
    // hits
    DevHit *hits = nullptr;
    unsigned long long *nhits = nullptr;
    long long cap = 1ll << 22;
    long long last_nhits = 0;
    int last_test = 'f';
    // tiles
    int4 *tiles = nullptr;
    size_t tiles_cap = 0;
    // stats
    double last_ms = 0.0;
    long long last_pairs = 0;
    int last_launches = 0;
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

#define HIPCHK(c, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((c), EPI_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                   \
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
        return fail(c, EPI_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
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
}

// ------------------------------------------------------------ kernels ----

__device__ __forceinline__ int decode(const uint32_t *w, int npack, int remain, int nid, int k)
{
    if (k >= nid) return -1;
    int j = k >> 4, t = k & 15;
    int sh = j < npack ? 30 - 2 * t : 2 * (remain - 1 - t);
    return (int)((w[j] >> sh) & 3u);
}

// One thread per (snp, 16-individual group): class masks for the MFMA form.
__global__ void transcode_planes(const uint32_t *words, int nid, int npack, int remain, int npack1,
                                 int nsnp, int Kp, uint8_t *planes, int32_t *miss)
{
    const int groups = Kp / 16;
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)groups * nsnp) return;
    int s = (int)(gid / groups), grp = (int)(gid % groups);
    const uint32_t *w = words + (size_t)s * npack1;
    uint32_t m[3][4] = {{0}};
    int anymiss = 0;
    for (int t = 0; t < 16; t++) {
        int c = decode(w, npack, remain, nid, 16 * grp + t);
        uint32_t byte = 0xFFu << (8 * (t & 3));
        if (c == 0) m[0][t >> 2] |= byte;
        if (c == 1) m[1][t >> 2] |= byte;
        if (c >= 0 && c != 3) m[2][t >> 2] |= byte;
        anymiss |= c == 3;
    }
    for (int p = 0; p < 3; p++)
        *reinterpret_cast<uint4 *>(planes + ((size_t)s * 3 + p) * Kp + 16 * grp) =
            make_uint4(m[p][0], m[p][1], m[p][2], m[p][3]);
    if (anymiss) atomicOr(&miss[s], 1);
}

// One thread per SNP: order the genotype classes by frequency. The plain form
// stores the two rarest classes as bitplanes 0 and 1 and derives the most
// common one; cls packs plane -> class (2 bits each).
__global__ void class_order(const uint32_t *words, int nid, int npack, int remain, int npack1, int nsnp,
                            int32_t *cls)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsnp) return;
    const uint32_t *w = words + (size_t)s * npack1;
    int n[4] = {0, 0, 0, 0};
    for (int k = 0; k < nid; k++) n[decode(w, npack, remain, nid, k)]++;
    int big = 0;
    if (n[1] > n[big]) big = 1;
    if (n[2] > n[big]) big = 2;
    int p0 = big == 0 ? 1 : 0;
    int p1 = big == 2 ? 1 : 2;
    cls[s] = p0 | (p1 << 2) | (big << 4);
}

// One thread per (word of 32 individuals, snp): class bitplanes, plain form.
__global__ void transcode_bits(const uint32_t *words, int nid, int npack, int remain, int npack1, int nsnp,
                               int Kw, const int32_t *cls, uint32_t *bits)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int wd = blockIdx.y;
    if (s >= nsnp) return;
    const uint32_t *w = words + (size_t)s * npack1;
    const int c0 = cls[s] & 3, c1 = (cls[s] >> 2) & 3;
    uint32_t b0 = 0, b1 = 0, bn = 0;
    for (int t = 0; t < 32; t++) {
        int c = decode(w, npack, remain, nid, 32 * wd + t);
        b0 |= (uint32_t)(c == c0) << t;
        b1 |= (uint32_t)(c == c1) << t;
        bn |= (uint32_t)(c >= 0 && c != 3) << t;
    }
    bits[((size_t)0 * Kw + wd) * nsnp + s] = b0;
    bits[((size_t)1 * Kw + wd) * nsnp + s] = b1;
    bits[((size_t)2 * Kw + wd) * nsnp + s] = bn;
}

// One workgroup per SNP: per genotype class (0, 1, 2) counts and phenotype
// sums over present phenotypes, for one phenotype; stored per class for the
// MFMA form and per bitplane (cls order) for the plain form.
__global__ __launch_bounds__(256) void marginals_kernel(const uint8_t *planes, int Kp, int nsnp,
                                                        const int8_t *wts, const int32_t *qv,
                                                        const float *yc, const int32_t *cls, Marg *marg)
{
    __shared__ long long sq[3][256];
    __shared__ double sy[3][256];
    __shared__ int sn[3][256];
    const int s = blockIdx.x, tid = threadIdx.x;
    int n[3] = {0, 0, 0};
    long long q[3] = {0, 0, 0};
    double y[3] = {0.0, 0.0, 0.0};
    const uint8_t *pl = planes + (size_t)s * 3 * Kp;
    for (int k = tid; k < Kp; k += 256) {
        if (!wts[k]) continue;  // row 0 = V
        if (!pl[(size_t)2 * Kp + k]) continue;  // missing call
        int c = pl[k] ? 0 : (pl[(size_t)Kp + k] ? 1 : 2);
        n[c] += 1;
        q[c] += qv[k];
        y[c] += (double)yc[k];
    }
    for (int p = 0; p < 3; p++) {
        sn[p][tid] = n[p];
        sq[p][tid] = q[p];
        sy[p][tid] = y[p];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int p = 0; p < 3; p++) {
                sn[p][tid] += sn[p][tid + st];
                sq[p][tid] += sq[p][tid + st];
                sy[p][tid] += sy[p][tid + st];
            }
        __syncthreads();
    }
    if (tid == 0) {
        Marg m;
        const int cl = cls[s];
        const int ntot = sn[0][0] + sn[1][0] + sn[2][0];
        m.n[0] = sn[0][0];
        m.n[1] = sn[1][0];
        m.n[2] = ntot;
        m.q[0] = sq[0][0];
        m.q[1] = sq[1][0];
        m.q[2] = sq[0][0] + sq[1][0] + sq[2][0];
        for (int p = 0; p < 2; p++) {
            int c = (cl >> (2 * p)) & 3;
            m.pn[p] = sn[c][0];
            m.ys[p] = sy[c][0];
        }
        m.pn[2] = ntot;
        m.ys[2] = sy[0][0] + sy[1][0] + sy[2][0];
        m.cls = cl;
        m.pad = 0;
        marg[s] = m;
    }
}

// ---------------------------------------------------------- host side ----

int ensure_marginals(epi_ctx *c)
{
    if (c->marg_ok) return EPI_OK;
    if (!c->planes || c->nphen == 0) return fail(c, EPI_ERR_STATE, "genotypes and phenotypes must be set");
    for (int r = 0; r < c->nphen; r++) {
        PhenDev &p = c->ph[r];
        int rc = dalloc(c, p.marg, (size_t)c->nsnp);
        if (rc) return rc;
        hipLaunchKernelGGL(marginals_kernel, dim3(c->nsnp), dim3(256), 0, c->stream, c->planes, c->Kp,
                           c->nsnp, p.wts, p.qv, p.yc, c->cls, p.marg);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->marg_ok = true;
    return EPI_OK;
}

struct TileLists {
    std::vector<int4> fast, gen;
    long long pairs = 0;
};

// Tiles of shape TA x TB over the block range, split into FAST tiles (no
// missing call in either SNP range) and GENERAL tiles of shape (TA/split) x TB.
// Within a block the tiles are emitted in super-tiles of 8 x 4 tiles so that
// neighbouring workgroups share SNP data in L2.
void build_tiles(const epi_ctx *c, int blk_begin, int blk_end, int TA, int TB, int split, TileLists &out)
{
    const int nchr = (int)c->chrstart.size();
    auto anymiss = [&](int s0, int s1) { return c->missc[s1] - c->missc[s0] > 0; };
    const int SA = 8, SB = 4;
    for (int i = 0; i < nchr; i++) {
        for (int j = 0; j <= i; j++) {
            int blk = i * (i + 1) / 2 + j;
            if (blk < blk_begin || blk >= blk_end) continue;
            int A0 = c->chrstart[i], A1 = A0 + c->chrsize[i];
            int B0 = c->chrstart[j], B1 = B0 + c->chrsize[j];
            if (A1 <= A0 || B1 <= B0) continue;
            // pairs in this block (a > b)
            for (long long a = A0; a < A1; a++) {
                long long lo = B0, hi = std::min<long long>(B1, a);
                if (hi > lo) out.pairs += hi - lo;
            }
            int nta = (A1 - A0 + TA - 1) / TA, ntb = (B1 - B0 + TB - 1) / TB;
            for (int sa = 0; sa < nta; sa += SA)
                for (int sb = 0; sb < ntb; sb += SB)
                    for (int ta = sa; ta < std::min(nta, sa + SA); ta++)
                        for (int tb = sb; tb < std::min(ntb, sb + SB); tb++) {
                            int a0 = A0 + ta * TA, a1 = std::min(A1, a0 + TA);
                            int b0 = B0 + tb * TB, b1 = std::min(B1, b0 + TB);
                            if (a1 - 1 <= b0) continue;  // no pair with a > b
                            if (!anymiss(a0, a1) && !anymiss(b0, b1)) {
                                out.fast.push_back(make_int4(a0, a1, b0, b1));
                            } else {
                                int sub = TA / split;
                                for (int x = a0; x < a1; x += sub) {
                                    int x1 = std::min(a1, x + sub);
                                    if (x1 - 1 <= b0) continue;
                                    out.gen.push_back(make_int4(x, x1, b0, b1));
                                }
                            }
                        }
        }
    }
}

int upload_tiles(epi_ctx *c, const TileLists &tl)
{
    size_t need = tl.fast.size() + tl.gen.size();
    if (need > c->tiles_cap) {
        int rc = dalloc(c, c->tiles, need);
        if (rc) return rc;
        c->tiles_cap = need;
    }
    if (!tl.fast.empty())
        HIPCHK(c, hipMemcpyAsync(c->tiles, tl.fast.data(), tl.fast.size() * sizeof(int4), hipMemcpyHostToDevice,
                                 c->stream));
    if (!tl.gen.empty())
        HIPCHK(c, hipMemcpyAsync(c->tiles + tl.fast.size(), tl.gen.data(), tl.gen.size() * sizeof(int4),
                                 hipMemcpyHostToDevice, c->stream));
    return EPI_OK;
}

HitSink make_sink(epi_ctx *c, const epi_scan_params *p)
{
    HitSink hs;
    hs.hits = c->hits;
    hs.nhits = c->nhits;
    hs.cap = c->cap;
    hs.test = p->test;
    hs.t1 = p->t1;
    hs.t2 = p->t2;
    hs.t1s = p->t1 * 0.98f;
    hs.t2s = p->t2 * 0.98f;
    hs.phen = p->phen;
    return hs;
}

int check_params(epi_ctx *c, const epi_scan_params *p)
{
    if (!p) return fail(c, EPI_ERR_ARG, "null params");
    if (p->test != 'f' && p->test != 'i') return fail(c, EPI_ERR_ARG, "test must be 'f' or 'i'");
    if (p->form != EPI_FORM_PLAIN && p->form != EPI_FORM_MFMA) return fail(c, EPI_ERR_ARG, "unknown form");
    if (p->phen < 0 || p->phen >= c->nphen) return fail(c, EPI_ERR_ARG, "phenotype index out of range");
    if (!(p->t1 == p->t1) || !(p->t2 == p->t2)) return fail(c, EPI_ERR_ARG, "threshold is NaN");
    if (!c->planes) return fail(c, EPI_ERR_STATE, "genotypes not set");
    if (c->chrstart.empty()) return fail(c, EPI_ERR_STATE, "chromosomes not set");
    return EPI_OK;
}

// Launch the pair kernels over the uploaded tile lists (fast tiles first).
template <bool DUMP>
int launch_pairs(epi_ctx *c, const epi_scan_params *p, const TileLists &tl, const DumpSink &ds)
{
    PhenDev &ph = c->ph[p->phen];
    const int nf = (int)tl.fast.size(), ng = (int)tl.gen.size();
    c->last_launches = 0;
    if (p->form == EPI_FORM_MFMA) {
        MfmaArgs A;
        A.planes = c->planes;
        A.wts = ph.wts;
        A.marg = ph.marg;
        A.nsnp = c->nsnp;
        A.Kp = c->Kp;
        A.nV = ph.nV;
        A.qV = ph.qV;
        A.sqV = ph.sqV;
        A.hs = make_sink(c, p);
        A.ds = ds;
        if (nf) {
            A.tiles = c->tiles;
            A.ntiles = nf;
            hipLaunchKernelGGL(mfma_fast_kernel<DUMP>, dim3(nf), dim3(256), 0, c->stream, A);
            HIPCHK(c, hipGetLastError());
            c->last_launches++;
        }
        if (ng) {
            A.tiles = c->tiles + nf;
            A.ntiles = ng;
            hipLaunchKernelGGL(mfma_general_kernel<DUMP>, dim3(ng), dim3(256), 0, c->stream, A);
            HIPCHK(c, hipGetLastError());
            c->last_launches++;
        }
    } else {
        PlainArgs P;
        P.bits = c->bits;
        P.yc = ph.yc;
        P.vbits = ph.vbits;
        P.marg = ph.marg;
        P.nsnp = c->nsnp;
        P.Kw = c->Kw;
        P.nV = ph.nV;
        P.ysV = ph.ysV;
        P.y2V = ph.y2V;
        P.hs = make_sink(c, p);
        P.ds = ds;
        if (nf) {
            P.tiles = c->tiles;
            P.ntiles = nf;
            hipLaunchKernelGGL((plain_kernel<2, 4, true, DUMP>), dim3(nf), dim3(256), 0, c->stream, P);
            HIPCHK(c, hipGetLastError());
            c->last_launches++;
        }
        if (ng) {
            P.tiles = c->tiles + nf;
            P.ntiles = ng;
            hipLaunchKernelGGL((plain_kernel<1, 4, false, DUMP>), dim3(ng), dim3(256), 0, c->stream, P);
            HIPCHK(c, hipGetLastError());
            c->last_launches++;
        }
    }
    return EPI_OK;
}

void tile_shape(int form, int &TA, int &TB, int &split)
{
    if (form == EPI_FORM_MFMA) {
        TA = 64;
        TB = 64;
        split = 4;
    } else {
        TA = 32;
        TB = 64;
        split = 2;
    }
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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return EPI_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        snprintf(c->name, sizeof(c->name), "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    }
    if (dalloc(c, c->nhits, 1) || dalloc(c, c->hits, (size_t)c->cap)) {
        epi_close(c);
        return EPI_ERR_NOMEM;
    }
    *out = c;
    return EPI_OK;
}

void epi_close(epi_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->words);
    dfree(c->planes);
    dfree(c->bits);
    dfree(c->dmiss);
    dfree(c->cls);
    for (auto &p : c->ph) free_phen(p);
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

int epi_set_genotypes(epi_ctx *c, const int32_t *words, int32_t nid, int32_t nsnp)
{
    if (!c) return EPI_ERR_ARG;
    if (!words || nid < 2 || nsnp < 2) return fail(c, EPI_ERR_ARG, "need nid >= 2, nsnp >= 2 and words");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nphen > 0 && nid != c->nid) {  // a new cohort: phenotypes of the old one no longer apply
        for (auto &p : c->ph) free_phen(p);
        c->ph.clear();
        c->nphen = 0;
    }
    c->nid = nid;
    c->nsnp = nsnp;
    c->npack = nid / 16;
    c->remain = nid % 16;
    c->npack1 = c->npack + (c->remain > 0);
    c->Kp = (nid + KPAD - 1) / KPAD * KPAD;
    c->Kw = c->Kp / 32;
    size_t nwords = (size_t)nsnp * c->npack1;
    int rc;
    if ((rc = dalloc(c, c->words, nwords)) || (rc = dalloc(c, c->planes, (size_t)nsnp * 3 * c->Kp)) ||
        (rc = dalloc(c, c->bits, (size_t)3 * c->Kw * nsnp)) || (rc = dalloc(c, c->dmiss, (size_t)nsnp)) ||
        (rc = dalloc(c, c->cls, (size_t)nsnp)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->words, words, nwords * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->dmiss, 0, (size_t)nsnp * sizeof(int32_t), c->stream));
    {
        long long total = (long long)(c->Kp / 16) * nsnp;
        int blocks = (int)((total + 255) / 256);
        hipLaunchKernelGGL(transcode_planes, dim3(blocks), dim3(256), 0, c->stream, (const uint32_t *)c->words, nid,
                           c->npack, c->remain, c->npack1, nsnp, c->Kp, c->planes, c->dmiss);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(class_order, dim3((nsnp + 255) / 256), dim3(256), 0, c->stream,
                           (const uint32_t *)c->words, nid, c->npack, c->remain, c->npack1, nsnp, c->cls);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(transcode_bits, dim3((nsnp + 255) / 256, c->Kw), dim3(256), 0, c->stream,
                           (const uint32_t *)c->words, nid, c->npack, c->remain, c->npack1, nsnp, c->Kw, c->cls,
                           c->bits);
        HIPCHK(c, hipGetLastError());
    }
    c->miss.assign(nsnp, 0);
    HIPCHK(c, hipMemcpyAsync(c->miss.data(), c->dmiss, (size_t)nsnp * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->missc.assign(nsnp + 1, 0);
    for (int s = 0; s < nsnp; s++) c->missc[s + 1] = c->missc[s] + (c->miss[s] != 0);
    c->marg_ok = false;
    // chromosome layout resets to one block over all SNPs
    c->chrstart.assign(1, 0);
    c->chrsize.assign(1, nsnp);
    c->snpchr.assign(nsnp, 0);
    c->last_nhits = 0;
    return EPI_OK;
}

int epi_set_phenotypes(epi_ctx *c, const float *phen, int32_t nphen, int32_t nid)
{
    if (!c) return EPI_ERR_ARG;
    if (!phen || nphen < 1 || nid < 2) return fail(c, EPI_ERR_ARG, "need phenotypes");
    if (c->nid && nid != c->nid) return fail(c, EPI_ERR_ARG, "nid %d differs from genotypes' nid %d", nid, c->nid);
    HIPCHK(c, hipSetDevice(c->device));
    const int Kp = (nid + KPAD - 1) / KPAD * KPAD, Kw = Kp / 32;
    for (auto &p : c->ph) free_phen(p);
    c->ph.assign(nphen, PhenDev());
    c->nphen = nphen;
    if (!c->nid) {
        c->nid = nid;
        c->Kp = Kp;
        c->Kw = Kw;
    }
    std::vector<int8_t> w((size_t)NW * Kp);
    std::vector<int32_t> qv(Kp);
    std::vector<float> yc(Kp);
    std::vector<uint32_t> vb(Kw);
    for (int r = 0; r < nphen; r++) {
        const float *y = phen + (size_t)r * nid;
        PhenDev &p = c->ph[r];
        double sum = 0.0;
        int nv = 0;
        for (int k = 0; k < nid; k++) {
            float v = y[k];
            if (v != v || v == -9.0f) continue;
            if (std::isinf(v)) return fail(c, EPI_ERR_RANGE, "phenotype %d has an infinite value", r);
            sum += v;
            nv++;
        }
        if (nv < 2) return fail(c, EPI_ERR_RANGE, "phenotype %d has fewer than 2 present values", r);
        double centre = sum / nv, maxabs = 0.0;
        for (int k = 0; k < nid; k++) {
            float v = y[k];
            if (v != v || v == -9.0f) continue;
            maxabs = std::max(maxabs, std::fabs((double)v - centre));
        }
        int f = 0;
        if (maxabs > 0.0) f = (int)std::floor(std::log2((double)QMAX / maxabs));
        while (maxabs > 0.0 && std::ldexp(maxabs, f) > (double)QMAX) f--;
        std::fill(w.begin(), w.end(), 0);
        std::fill(qv.begin(), qv.end(), 0);
        std::fill(yc.begin(), yc.end(), 0.0f);
        std::fill(vb.begin(), vb.end(), 0u);
        long long qsum = 0;
        double sq = 0.0, ys = 0.0, y2 = 0.0;
        for (int k = 0; k < nid; k++) {
            float v = y[k];
            if (v != v || v == -9.0f) continue;
            double d = (double)v - centre;
            long long q = std::llround(std::ldexp(d, f));
            if (q > QMAX || q < -QMAX) return fail(c, EPI_ERR_RANGE, "fixed-point phenotype out of range");
            long long l0 = ((q % 256) + 256 + 128) % 256 - 128;
            long long q1 = (q - l0) / 256;
            long long l1 = ((q1 % 256) + 256 + 128) % 256 - 128;
            long long l2 = (q1 - l1) / 256;
            long long z = std::llround((double)q * (double)q / ZSHIFT);
            long long z0 = ((z % 256) + 256 + 128) % 256 - 128;
            long long z1q = (z - z0) / 256;
            long long z1 = ((z1q % 256) + 256 + 128) % 256 - 128;
            long long z2 = (z1q - z1) / 256;
            if (l2 < -128 || l2 > 127 || z2 < -128 || z2 > 127) return fail(c, EPI_ERR_RANGE, "limb overflow");
            w[0 * (size_t)Kp + k] = 1;
            w[1 * (size_t)Kp + k] = (int8_t)l0;
            w[2 * (size_t)Kp + k] = (int8_t)l1;
            w[3 * (size_t)Kp + k] = (int8_t)l2;
            w[4 * (size_t)Kp + k] = (int8_t)z0;
            w[5 * (size_t)Kp + k] = (int8_t)z1;
            w[6 * (size_t)Kp + k] = (int8_t)z2;
            qv[k] = (int32_t)q;
            qsum += q;
            sq += (double)q * (double)q;
            float ycf = (float)d;
            yc[k] = ycf;
            ys += (double)ycf;
            y2 += (double)ycf * (double)ycf;
            vb[k >> 5] |= 1u << (k & 31);
        }
        p.fbits = f;
        p.centre = centre;
        p.nV = nv;
        p.qV = qsum;
        p.sqV = sq;
        p.ysV = ys;
        p.y2V = y2;
        int rc;
        if ((rc = dalloc(c, p.wts, w.size())) || (rc = dalloc(c, p.qv, (size_t)Kp)) ||
            (rc = dalloc(c, p.yc, (size_t)Kp)) || (rc = dalloc(c, p.vbits, (size_t)Kw)))
            return rc;
        HIPCHK(c, hipMemcpyAsync(p.wts, w.data(), w.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(p.qv, qv.data(), (size_t)Kp * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(p.yc, yc.data(), (size_t)Kp * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(p.vbits, vb.data(), (size_t)Kw * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->marg_ok = false;
    return EPI_OK;
}

int epi_set_chromosomes(epi_ctx *c, const int32_t *chrstart, const int32_t *chrsize, int32_t nchr)
{
    if (!c) return EPI_ERR_ARG;
    if (!chrstart || !chrsize || nchr < 1) return fail(c, EPI_ERR_ARG, "need chromosomes");
    if (!c->nsnp) return fail(c, EPI_ERR_STATE, "set genotypes before chromosomes");
    std::vector<int> snpchr(c->nsnp, -1);
    for (int i = 0; i < nchr; i++) {
        if (chrsize[i] < 0 || chrstart[i] < 0 || (long long)chrstart[i] + chrsize[i] > c->nsnp)
            return fail(c, EPI_ERR_ARG, "chromosome %d out of range", i);
        for (int s = chrstart[i]; s < chrstart[i] + chrsize[i]; s++) {
            if (snpchr[s] >= 0) return fail(c, EPI_ERR_ARG, "chromosomes %d and %d overlap", snpchr[s], i);
            snpchr[s] = i;
        }
    }
    c->chrstart.assign(chrstart, chrstart + nchr);
    c->chrsize.assign(chrsize, chrsize + nchr);
    c->snpchr.swap(snpchr);
    return EPI_OK;
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

int epi_scan(epi_ctx *c, const epi_scan_params *p)
{
    if (!c) return EPI_ERR_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ensure_marginals(c))) return rc;
    int nchr = (int)c->chrstart.size();
    int nblk = nchr * (nchr + 1) / 2;
    int bb = std::max(0, p->blk_begin), be = p->blk_end <= 0 ? nblk : std::min(nblk, p->blk_end);
    int TA, TB, split;
    tile_shape(p->form, TA, TB, split);
    TileLists tl;
    build_tiles(c, bb, be, TA, TB, split, tl);
    if ((rc = upload_tiles(c, tl))) return rc;
    HIPCHK(c, hipMemsetAsync(c->nhits, 0, sizeof(unsigned long long), c->stream));
    DumpSink ds;
    memset(&ds, 0, sizeof(ds));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if ((rc = launch_pairs<false>(c, p, tl, ds))) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long nh = 0;
    HIPCHK(c, hipMemcpyAsync(&nh, c->nhits, sizeof(nh), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_pairs = tl.pairs;
    c->last_nhits = (long long)nh;
    c->last_test = p->test;
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
    auto blk_of = [&](const DevHit &x) {
        int i = c->snpchr[x.snp1], j = c->snpchr[x.snp2];
        return i * (i + 1) / 2 + j;
    };
    std::vector<long long> key((size_t)have);
    for (long long t = 0; t < have; t++) key[t] = t;
    std::sort(key.begin(), key.end(), [&](long long x, long long y) {
        const DevHit &p = h[x], &q = h[y];
        int bp = blk_of(p), bq = blk_of(q);
        if (bp != bq) return bp < bq;
        if (p.snp1 != q.snp1) return p.snp1 < q.snp1;
        if (p.snp2 != q.snp2) return p.snp2 < q.snp2;
        return p.phen < q.phen;
    });
    long long m = std::min<long long>(have, cap);
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

int epi_pair_dump(epi_ctx *c, const epi_scan_params *p, int32_t a0, int32_t a1, int32_t b0, int32_t b1,
                  int32_t *counts, double *sums, double *F, double *F2)
{
    if (!c) return EPI_ERR_ARG;
    int rc = check_params(c, p);
    if (rc) return rc;
    if (a0 < 0 || b0 < 0 || a1 > c->nsnp || b1 > c->nsnp || a1 <= a0 || b1 <= b0)
        return fail(c, EPI_ERR_ARG, "dump rectangle out of range");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ensure_marginals(c))) return rc;
    int TA, TB, split;
    tile_shape(p->form, TA, TB, split);
    // Tiles over the rectangle in global SNP space (no block restriction).
    TileLists all;
    {
        // emit tiles covering [a0,a1) x [b0,b1) with the same FAST/GENERAL split
        auto anymiss = [&](int s0, int s1) { return c->missc[s1] - c->missc[s0] > 0; };
        for (int x = a0; x < a1; x += TA)
            for (int y = b0; y < b1; y += TB) {
                int x1 = std::min(a1, x + TA), y1 = std::min(b1, y + TB);
                if (x1 - 1 <= y) continue;
                if (!anymiss(x, x1) && !anymiss(y, y1)) {
                    all.fast.push_back(make_int4(x, x1, y, y1));
                } else {
                    int sub = TA / split;
                    for (int u = x; u < x1; u += sub) {
                        int u1 = std::min(x1, u + sub);
                        if (u1 - 1 <= y) continue;
                        all.gen.push_back(make_int4(u, u1, y, y1));
                    }
                }
            }
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
    PhenDev &ph = c->ph[p->phen];
    ds.unit = p->form == EPI_FORM_MFMA ? std::ldexp(1.0, -ph.fbits) : 1.0;
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
    if (hrc == EPI_OK) hrc = launch_pairs<true>(c, p, all, ds);
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

}  // extern "C"
