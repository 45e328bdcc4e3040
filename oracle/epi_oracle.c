/*
 * ORACLE - test infrastructure only. Nothing in the product links, loads or
 * calls this file; the tests, the smoke check and the benchmark's CPU-baseline
 * leg use it as the checker.
 *
 * FP64 CPU restatement of the per-pair semantics of epiGPU's pair-scan kernels
 * `kernelfull` (src/EpiOpenCL.c:45-264) and `kernelint` (src/EpiOpenCL.c:276-504)
 * and of the pair/block order of perform_scan (src/EpiOpenCL.c:1527-1607).
 *
 * Per pair (a, b), a on the outer chromosome, global index a > b:
 *   complete cases I = { k : g_a(k) != 3, g_b(k) != 3, y_k not missing }
 *                                          (src/EpiOpenCL.c:135, :362)
 *   missing phenotype: y == -9 (the reference's sentinel) or NaN
 *   cell c = 3 g_a + g_b, n_c, S_c = sum of y over the cell (:140-142)
 *   nfac = #{c : n_c > 0}, d1 = nfac - 1, d2 = N' - nfac        (:231-247)
 *   SSB = sum_{n_c>0} n_c (m_c - mbar)^2                       (:235-243)
 *   SSW = sum_I (y - m_c(k))^2  (= SST - SSB of the reference, two-pass here)
 *   F   = (SSB / d1) / (SSW / d2)                               (:245-247)
 *   R_r = row mean over g_a = r, C_s = column mean over g_b = s (:463-468)
 *   SSI = sum_{n_c>0} n_c (m_c - R_{g_a} - C_{g_b} + mbar)^2    (:469-481)
 *   F2  = (SSI / 4) / (SSW / d2)                                (:482-486)
 *   hit, test 'f': F > t1                                       (:251)
 *   hit, test 'i': (F > t1 || F2 > t2) && nfac > 6              (:487-489)
 * F is only defined for d1 >= 1 and d2 >= 1; other pairs are never hits.
 *
 * Safe mode (-s 1; kernels kernelfullsafe / kernelintsafe, src/EpiOpenCL.c:505-963,
 * selected at :1486-1499): complete cases only require both calls, so -9 is an
 * ordinary phenotype value (:597-598, no sentinel test), and test 'i' has no
 * nfac > 6 condition (:949-950). NaN stays missing here in both modes.
 *
 * Everything is computed from the raw genotypes and FP64 copies of the FP32
 * phenotype, two-pass, so the only rounding is FP64's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_hit {
    int32_t snp1, snp2, df1, df2;
    double F, F2;
    int32_t block;
    int32_t pad;
} orc_hit;

static inline int y_missing(float y, int safe) { return y != y || (!safe && y == -9.0f); }

/* Decode layout A1 (src/readplink.c:1104-1143). */
void orc_unpack(const int32_t *words, int nid, int nsnp, int8_t *codes)
{
    int npack = nid / 16, remain = nid % 16, npack1 = npack + (remain > 0);
    for (int s = 0; s < nsnp; s++) {
        const uint32_t *w = (const uint32_t *)words + (size_t)s * npack1;
        int8_t *g = codes + (size_t)s * nid;
        for (int k = 0; k < nid; k++) {
            int j = k / 16, t = k % 16;
            int sh = j < npack ? 30 - 2 * t : 2 * (remain - 1 - t);
            g[k] = (int8_t)((w[j] >> sh) & 3u);
        }
    }
}

/* Cell counts and FP64 cell sums over complete cases. */
void orc_pair_table(const int8_t *ga, const int8_t *gb, const float *y, int nid, int safe,
                    int64_t counts[9], double sums[9])
{
    for (int c = 0; c < 9; c++) {
        counts[c] = 0;
        sums[c] = 0.0;
    }
    for (int k = 0; k < nid; k++) {
        if (ga[k] == 3 || gb[k] == 3 || y_missing(y[k], safe)) continue;
        int c = 3 * ga[k] + gb[k];
        counts[c]++;
        sums[c] += (double)y[k];
    }
}

/* Full per-pair statistics (two-pass). Returns nfac. */
int orc_pair_stats(const int8_t *ga, const int8_t *gb, const float *y, int nid, int safe,
                   int64_t counts[9], double sums[9], double *F, double *F2,
                   int *d1, int *d2)
{
    orc_pair_table(ga, gb, y, nid, safe, counts, sums);
    int64_t n = 0;
    double S = 0.0;
    int nfac = 0;
    for (int c = 0; c < 9; c++) {
        n += counts[c];
        S += sums[c];
        nfac += counts[c] > 0;
    }
    *d1 = nfac - 1;
    *d2 = (int)(n - nfac);
    if (n == 0) {
        *F = NAN;
        *F2 = NAN;
        return nfac;
    }
    double mbar = S / (double)n, m[9];
    for (int c = 0; c < 9; c++) m[c] = counts[c] ? sums[c] / (double)counts[c] : 0.0;
    double ssw = 0.0;
    for (int k = 0; k < nid; k++) {
        if (ga[k] == 3 || gb[k] == 3 || y_missing(y[k], safe)) continue;
        double e = (double)y[k] - m[3 * ga[k] + gb[k]];
        ssw += e * e;
    }
    double ssb = 0.0, ssi = 0.0, R[3], C[3];
    for (int r = 0; r < 3; r++) {
        double sr = sums[3 * r] + sums[3 * r + 1] + sums[3 * r + 2];
        double nr = (double)(counts[3 * r] + counts[3 * r + 1] + counts[3 * r + 2]);
        double sc = sums[r] + sums[r + 3] + sums[r + 6];
        double nc = (double)(counts[r] + counts[r + 3] + counts[r + 6]);
        R[r] = nr > 0 ? sr / nr : 0.0;
        C[r] = nc > 0 ? sc / nc : 0.0;
    }
    for (int c = 0; c < 9; c++) {
        if (!counts[c]) continue;
        double e = m[c] - mbar;
        ssb += (double)counts[c] * e * e;
        double t = m[c] - R[c / 3] - C[c % 3] + mbar;
        ssi += (double)counts[c] * t * t;
    }
    double msw = ssw / (double)(*d2);
    *F = (ssb / (double)(*d1)) / msw;
    *F2 = (ssi / 4.0) / msw;
    return nfac;
}

static inline int is_hit(int test, int safe, double F, double F2, int nfac, int d1, int d2, float t1, float t2)
{
    if (d1 < 1 || d2 < 1) return 0;
    if (test == 'f') return F > (double)t1;
    return (F > (double)t1 || F2 > (double)t2) && (safe || nfac > 6);
}

/* Per-pair results for an explicit list of (a, b) pairs. */
void orc_pairs(const int8_t *codes, int nid, const float *y, int safe, const int32_t *pairs,
               int64_t npairs, int32_t *counts_out, double *sums_out, double *F,
               double *F2, int32_t *d1, int32_t *d2, int32_t *nfac)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < npairs; p++) {
        int64_t cnt[9];
        double sm[9], f, f2;
        int a = pairs[2 * p], b = pairs[2 * p + 1], e1, e2;
        int nf = orc_pair_stats(codes + (size_t)a * nid, codes + (size_t)b * nid, y, nid, safe,
                                cnt, sm, &f, &f2, &e1, &e2);
        for (int c = 0; c < 9; c++) {
            if (counts_out) counts_out[9 * p + c] = (int32_t)cnt[c];
            if (sums_out) sums_out[9 * p + c] = sm[c];
        }
        F[p] = f;
        F2[p] = f2;
        if (d1) d1[p] = e1;
        if (d2) d2[p] = e2;
        if (nfac) nfac[p] = nf;
    }
}

/*
 * Exhaustive scan over chr x chr blocks [blk_begin, blk_end) in the
 * reference order (block index i(i+1)/2 + j for j <= i, src/EpiOpenCL.c:1527,
 * :1540, :1445-1456). Pairs a in chromosome i, b in chromosome j, a > b.
 * Returns the number of hits (all of them, also beyond cap).
 */
int64_t orc_scan(const int8_t *codes, int nid, const float *y, int safe, const int32_t *chrstart,
                 const int32_t *chrsize, int nchr, int test, float t1, float t2,
                 int blk_begin, int blk_end, orc_hit *hits, int64_t cap)
{
    int64_t nh = 0;
    for (int i = 0; i < nchr; i++) {
        for (int j = 0; j <= i; j++) {
            int blk = i * (i + 1) / 2 + j;
            if (blk < blk_begin || blk >= blk_end) continue;
            int a0 = chrstart[i], a1 = a0 + chrsize[i];
            int b0 = chrstart[j], b1 = b0 + chrsize[j];
#pragma omp parallel for schedule(dynamic, 1)
            for (int a = a0; a < a1; a++) {
                for (int b = b0; b < b1 && b < a; b++) {
                    int64_t cnt[9];
                    double sm[9], f, f2;
                    int e1, e2;
                    int nf = orc_pair_stats(codes + (size_t)a * nid, codes + (size_t)b * nid, y,
                                            nid, safe, cnt, sm, &f, &f2, &e1, &e2);
                    if (!is_hit(test, safe, f, f2, nf, e1, e2, t1, t2)) continue;
                    int64_t slot;
#pragma omp atomic capture
                    slot = nh++;
                    if (slot < cap) {
                        orc_hit *h = &hits[slot];
                        h->snp1 = a;
                        h->snp2 = b;
                        h->df1 = e1;
                        h->df2 = e2;
                        h->F = f;
                        h->F2 = test == 'f' ? 0.0 : f2;
                        h->block = blk;
                        h->pad = 0;
                    }
                }
            }
        }
    }
    return nh;
}
