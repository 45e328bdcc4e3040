/*
 * ORACLE - test infrastructure only (CPU baseline leg of the benchmark and
 * error comparisons in the tests). Nothing in the product links or calls it.
 *
 * FP32 CPU restatement of the *shape* of epiGPU's normal-mode kernels, used as
 * the CPU baseline because the reference OpenCL scan cannot run on a CPU here
 * (it enumerates CL_DEVICE_TYPE_GPU only, src/EpiOpenCL.c:1394, :1474).
 * Per pair it does what one work-item of `square` does:
 *   - phenotype mean and SS over all nid in FP32          src/EpiOpenCL.c:1364-1376
 *   - per individual: 2-bit decode of both SNPs' words    :347-360, :423-450
 *   - complete case: uint16 count and FP32 sum of the cell :362-366
 *   - otherwise remove y from mean/SS by reverse Welford   :367-370
 *   - FP32 epilogue: nfac, SSB, (row/col means, SSI), F, F2 :455-486
 *   - hit rule 'f': F > t1; 'i': (F > t1 || F2 > t2) && nfac > 6
 * Missing phenotypes are -9 as in the reference; NaN is mapped to -9 first.
 * OpenMP over pairs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct ref32_out {
    float F, F2;
    int32_t df1, df2, nfac, hit;
} ref32_out;

static inline int code_of(const uint32_t *w, int npack, int remain, int k)
{
    int j = k >> 4, t = k & 15;
    int sh = j < npack ? 30 - 2 * t : 2 * (remain - 1 - t);
    return (int)((w[j] >> sh) & 3u);
}

void ref32_phen_stats(const float *phen, int nid, float *mY, float *SSY)
{
    float m = 0.0f, ss = 0.0f;
    for (int i = 0; i < nid; i++) m += phen[i];
    m /= (float)nid;
    for (int i = 0; i < nid; i++) ss += (phen[i] - m) * (phen[i] - m);
    *mY = m;
    *SSY = ss;
}

static void ref32_pair(const uint32_t *wa, const uint32_t *wb, const float *phen, int nid,
                       float mY, float SSY, int test, float t1, float t2, ref32_out *o)
{
    int npack = nid / 16, remain = nid % 16;
    unsigned short cnt[9] = {0};
    float sm[9] = {0};
    float locmy = mY, locssy = SSY;
    int locnid = nid;
    for (int k = 0; k < nid; k++) {
        int g1 = code_of(wa, npack, remain, k), g2 = code_of(wb, npack, remain, k);
        float y = phen[k];
        if (g1 != 3 && g2 != 3 && y != -9.0f) {
            int c = 3 * g1 + g2;
            cnt[c] += 1;
            sm[c] += y;
        } else {
            float delta = (locnid * locmy - y) / (float)(--locnid);
            locssy -= (y - locmy) * (y - delta);
            locmy = delta;
        }
    }
    int nfac = 0;
    for (int c = 0; c < 9; c++) nfac += cnt[c] > 0;
    float ssb = 0.0f, ssi = 0.0f, mr[3], mc[3];
    for (int i = 0; i < 3; i++) {
        mc[i] = (sm[i] + sm[i + 3] + sm[i + 6]) / (float)(cnt[i] + cnt[i + 3] + cnt[i + 6]);
        mr[i] = (sm[3 * i] + sm[3 * i + 1] + sm[3 * i + 2]) /
                (float)(cnt[3 * i] + cnt[3 * i + 1] + cnt[3 * i + 2]);
    }
    for (int c = 0; c < 9; c++) {
        if (!cnt[c]) continue;
        float m = sm[c] / (float)cnt[c];
        ssb += cnt[c] * ((m - locmy) * (m - locmy));
        float t = m - mr[c / 3] - mc[c % 3] + locmy;
        ssi += cnt[c] * t * t;
    }
    int d1 = nfac - 1, d2 = locnid - nfac;
    float msw = (locssy - ssb) / (float)d2;
    float F = (ssb / (float)d1) / msw, F2 = (ssi / 4.0f) / msw;
    o->F = F;
    o->F2 = test == 'f' ? 0.0f : F2;
    o->df1 = d1;
    o->df2 = d2;
    o->nfac = nfac;
    o->hit = test == 'f' ? (F > t1) : ((F > t1 || F2 > t2) && nfac > 6);
}

/* Map NaN phenotypes to the reference's -9 sentinel (copy). */
static float *phen_ref(const float *phen, int nid)
{
    float *p = (float *)malloc(sizeof(float) * (size_t)nid);
    for (int i = 0; i < nid; i++) p[i] = phen[i] != phen[i] ? -9.0f : phen[i];
    return p;
}

/* Per-pair results for a list of pairs. */
void ref32_pairs(const int32_t *words, int nid, const float *phen, const int32_t *pairs,
                 int64_t npairs, int test, float t1, float t2, ref32_out *out)
{
    int npack1 = nid / 16 + (nid % 16 > 0);
    float *p = phen_ref(phen, nid), mY, SSY;
    ref32_phen_stats(p, nid, &mY, &SSY);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < npairs; i++) {
        const uint32_t *wa = (const uint32_t *)words + (size_t)pairs[2 * i] * npack1;
        const uint32_t *wb = (const uint32_t *)words + (size_t)pairs[2 * i + 1] * npack1;
        ref32_pair(wa, wb, p, nid, mY, SSY, test, t1, t2, &out[i]);
    }
    free(p);
}

/*
 * Scan of square tiles: tile t covers a in [tiles[2t], +tile), b in
 * [tiles[2t+1], +tile), pairs with a > b and both < nsnp. Returns the pairs
 * tested; *nhits gets the hit count. This is the timed CPU-baseline kernel.
 */
int64_t ref32_scan_tiles(const int32_t *words, int nid, int nsnp, const float *phen,
                         const int32_t *tiles, int ntiles, int tile, int test, float t1,
                         float t2, int64_t *nhits)
{
    int npack1 = nid / 16 + (nid % 16 > 0);
    float *p = phen_ref(phen, nid), mY, SSY;
    ref32_phen_stats(p, nid, &mY, &SSY);
    int64_t pairs = 0, hits = 0;
#pragma omp parallel for schedule(dynamic, 1) collapse(2) reduction(+ : pairs, hits)
    for (int t = 0; t < ntiles; t++) {
        for (int da = 0; da < tile; da++) {
            int a = tiles[2 * t] + da;
            if (a >= nsnp) continue;
            for (int b = tiles[2 * t + 1]; b < tiles[2 * t + 1] + tile && b < a; b++) {
                ref32_out o;
                ref32_pair((const uint32_t *)words + (size_t)a * npack1,
                           (const uint32_t *)words + (size_t)b * npack1, p, nid, mY, SSY, test,
                           t1, t2, &o);
                pairs++;
                hits += o.hit;
            }
        }
    }
    free(p);
    *nhits = hits;
    return pairs;
}
