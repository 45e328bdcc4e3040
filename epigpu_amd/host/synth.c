/*
 * Seeded synthetic workload generator for the scan.
 *
 * Every SNP and every phenotype owns an independent xoshiro256** stream
 * (Blackman & Vigna, 2018) whose 256-bit state is filled from splitmix64
 * started at a hash of (seed, purpose, id). The draw order is fixed and
 * documented in epihost.h, so a dataset is a pure function of its parameters
 * and generation parallelises over SNPs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "epihost.h"

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_init(rng_t *r, uint64_t seed, uint64_t purpose, uint64_t id)
{
    uint64_t h = seed;
    uint64_t k = splitmix64(&h) ^ (purpose * 0xD6E8FEB86659FD93ull);
    k ^= splitmix64(&k) + id * 0xA0761D6478BD642Full;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&k);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t out = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}

/* Uniform on [0,1) with 53 random bits. */
static inline double rng_u(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
/* Uniform on (0,1), safe for log. */
static inline double rng_u_open(rng_t *r) { return ((double)(rng_next(r) >> 11) + 0.5) * 0x1.0p-53; }

static inline double rng_normal(rng_t *r)
{
    double u1 = rng_u_open(r), u2 = rng_u(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

static void pack_one(const int8_t *g, int nid, int32_t *w)
{
    int npack = nid / 16, remain = nid % 16;
    for (int j = 0; j < npack; j++) {
        uint32_t v = 0;
        for (int t = 0; t < 16; t++) v |= (uint32_t)(g[16 * j + t] & 3) << (30 - 2 * t);
        w[j] = (int32_t)v;
    }
    if (remain) {
        uint32_t v = 0;
        for (int t = 0; t < remain; t++) v |= (uint32_t)(g[16 * npack + t] & 3) << (2 * (remain - 1 - t));
        w[npack] = (int32_t)v;
    }
}

static inline int code_at(const int32_t *w, int nid, int k)
{
    int npack = nid / 16, remain = nid % 16;
    int j = k / 16, t = k % 16;
    uint32_t v = (uint32_t)w[j];
    if (j < npack) return (int)((v >> (30 - 2 * t)) & 3);
    return (int)((v >> (2 * (remain - 1 - t))) & 3);
}

void epi_synth_genotypes(uint64_t seed, int nid, int nsnp, double maf_lo,
                         double maf_hi, double miss, int32_t *words)
{
    int npack1 = nid / 16 + (nid % 16 > 0);
#pragma omp parallel
    {
        int8_t *g = (int8_t *)malloc((size_t)nid);
#pragma omp for schedule(dynamic, 64)
        for (int s = 0; s < nsnp; s++) {
            rng_t r;
            rng_init(&r, seed, 1, (uint64_t)s);
            double maf = maf_lo + (maf_hi - maf_lo) * rng_u(&r);
            for (int k = 0; k < nid; k++) {
                int a = rng_u(&r) < maf;
                int b = rng_u(&r) < maf;
                int c = a + b;
                if (miss > 0.0 && rng_u(&r) < miss) c = EPI_MISS;
                g[k] = (int8_t)c;
            }
            pack_one(g, nid, words + (size_t)s * npack1);
        }
        free(g);
    }
}

void epi_synth_phenotype(uint64_t seed, int r_id, int nid, int dist, double miss, float *phen)
{
    rng_t r;
    rng_init(&r, seed, 2, (uint64_t)r_id);
    for (int k = 0; k < nid; k++) {
        double v;
        switch (dist) {
        case EPI_PHEN_GAMMA2:
            v = -log(rng_u_open(&r)) - log(rng_u_open(&r));
            break;
        case EPI_PHEN_GAMMA2_STD:
            v = (-log(rng_u_open(&r)) - log(rng_u_open(&r)) - 2.0) / sqrt(2.0);
            break;
        case EPI_PHEN_T5: {
            double z = rng_normal(&r), c = 0.0;
            for (int i = 0; i < 5; i++) {
                double e = rng_normal(&r);
                c += e * e;
            }
            v = z / sqrt(c / 5.0);
            break;
        }
        case EPI_PHEN_BERN02:
            v = rng_u(&r) < 0.2 ? 1.0 : 0.0;
            break;
        default:
            v = rng_normal(&r);
            break;
        }
        if (miss > 0.0 && rng_u(&r) < miss) v = NAN;
        phen[k] = (float)v;
    }
}

void epi_synth_plant(uint64_t seed, const int32_t *words, int nid, int nsnp,
                     float *phen, int npairs, float effect, int32_t *pairs_out)
{
    rng_t r;
    int npack1 = nid / 16 + (nid % 16 > 0);
    rng_init(&r, seed, 3, 0);
    for (int p = 0; p < npairs; p++) {
        int a, b;
        do {
            a = (int)(rng_u(&r) * nsnp);
            b = (int)(rng_u(&r) * nsnp);
        } while (a == b);
        if (a < b) { int t = a; a = b; b = t; }
        int ga = (int)(rng_u(&r) * 3), gb = (int)(rng_u(&r) * 3);
        const int32_t *wa = words + (size_t)a * npack1, *wb = words + (size_t)b * npack1;
        for (int k = 0; k < nid; k++) {
            float y = phen[k];
            if (y != y || y == -9.0f) continue;
            if (code_at(wa, nid, k) == ga && code_at(wb, nid, k) == gb) phen[k] = y + effect;
        }
        if (pairs_out) {
            pairs_out[4 * p + 0] = a;
            pairs_out[4 * p + 1] = b;
            pairs_out[4 * p + 2] = ga;
            pairs_out[4 * p + 3] = gb;
        }
    }
}

void epi_synth_chromosomes(int nchr, const int *sizes, chromosome *chrstat)
{
    int start = 0;
    for (int i = 0; i < nchr; i++) {
        memset(&chrstat[i], 0, sizeof(chromosome));
        chrstat[i].chrid = i + 1;
        snprintf(chrstat[i].chrname, sizeof(chrstat[i].chrname), "%d", i + 1);
        chrstat[i].chrsize = sizes[i];
        chrstat[i].chrstart = start;
        start += sizes[i];
    }
}
