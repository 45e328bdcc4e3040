/*
 * 2-bit genotype codec, layout A1 of the reference (src/readplink.c:1066-1143).
 *
 * Words are SNP-major; npack = nid / 16 full words hold 16 genotypes each with
 * the first individual in the two most significant bits; when remain = nid % 16
 * is non-zero a tail word holds the last `remain` genotypes right-aligned, the
 * first of them in bits [2*remain-1 : 2*remain-2].
 */
#include "epihost.h"

void epi_layout(int nid, int *npack, int *remain, int *npack1)
{
    int np = nid / 16, r = nid % 16;
    if (npack) *npack = np;
    if (remain) *remain = r;
    if (npack1) *npack1 = np + (r > 0);
}

void epi_pack(const int8_t *codes, int nid, int nsnp, int32_t *words)
{
    int npack, remain, npack1;
    epi_layout(nid, &npack, &remain, &npack1);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsnp; s++) {
        const int8_t *g = codes + (size_t)s * nid;
        int32_t *w = words + (size_t)s * npack1;
        for (int j = 0; j < npack; j++) {
            uint32_t v = 0;
            for (int t = 0; t < 16; t++)
                v |= (uint32_t)(g[16 * j + t] & 3) << (30 - 2 * t);
            w[j] = (int32_t)v;
        }
        if (remain) {
            uint32_t v = 0;
            for (int t = 0; t < remain; t++)
                v |= (uint32_t)(g[16 * npack + t] & 3) << (2 * (remain - 1 - t));
            w[npack] = (int32_t)v;
        }
    }
}

void epi_unpack(const int32_t *words, int nid, int nsnp, int8_t *codes)
{
    int npack, remain, npack1;
    epi_layout(nid, &npack, &remain, &npack1);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < nsnp; s++) {
        int8_t *g = codes + (size_t)s * nid;
        const int32_t *w = words + (size_t)s * npack1;
        for (int j = 0; j < npack; j++) {
            uint32_t v = (uint32_t)w[j];
            for (int t = 0; t < 16; t++) g[16 * j + t] = (int8_t)((v >> (30 - 2 * t)) & 3);
        }
        if (remain) {
            uint32_t v = (uint32_t)w[npack];
            for (int t = 0; t < remain; t++)
                g[16 * npack + t] = (int8_t)((v >> (2 * (remain - 1 - t))) & 3);
        }
    }
}
