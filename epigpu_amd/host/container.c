/*
 * epiGPU packed binary container (src/readplink.c:1145-1273).
 *
 * Byte layout, native endianness (x86-64 little endian, gcc record padding):
 *   6 bytes "epiGPU" (no terminator)
 *   int32 nid, nsnp, npack, remain, nchr
 *   ped[nid]          128 bytes each
 *   map[nsnp]          88 bytes each
 *   chromosome[nchr]   44 bytes each
 *   int32 words[npack1 * nsnp]   genotype layout A1
 * The reader validates the signature and that npack/remain agree with nid,
 * and reports truncation instead of reading garbage.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "epihost.h"

_Static_assert(sizeof(ped) == 128, "ped record must match the container layout");
_Static_assert(sizeof(map) == 88, "map record must match the container layout");
_Static_assert(sizeof(chromosome) == 44, "chromosome record must match the container layout");

static const char SIG[6] = {'e', 'p', 'i', 'G', 'P', 'U'};

void epi_container_free(epi_container *c)
{
    if (!c) return;
    free(c->dat);
    free(c->genmap);
    free(c->chrstat);
    free(c->geno);
    c->dat = NULL;
    c->genmap = NULL;
    c->chrstat = NULL;
    c->geno = NULL;
}

int epi_read_container(const char *path, epi_container *c)
{
    memset(c, 0, sizeof(*c));
    FILE *f = fopen(path, "rb");
    if (!f) return EPIH_ERR_IO;
    char sig[6];
    int hdr[5];
    if (fread(sig, 1, 6, f) != 6 || memcmp(sig, SIG, 6) != 0) {
        fclose(f);
        return EPIH_ERR_FORMAT;
    }
    if (fread(hdr, sizeof(int), 5, f) != 5) {
        fclose(f);
        return EPIH_ERR_FORMAT;
    }
    c->nid = hdr[0];
    c->nsnp = hdr[1];
    c->npack = hdr[2];
    c->remain = hdr[3];
    c->nchr = hdr[4];
    if (c->nid <= 0 || c->nsnp <= 0 || c->nchr <= 0 || c->npack != c->nid / 16 ||
        c->remain != c->nid % 16) {
        fclose(f);
        return EPIH_ERR_FORMAT;
    }
    size_t npack1 = (size_t)c->npack + (c->remain > 0);
    c->dat = (ped *)malloc(sizeof(ped) * (size_t)c->nid);
    c->genmap = (map *)malloc(sizeof(map) * (size_t)c->nsnp);
    c->chrstat = (chromosome *)malloc(sizeof(chromosome) * (size_t)c->nchr);
    c->geno = (int32_t *)malloc(sizeof(int32_t) * npack1 * (size_t)c->nsnp);
    if (!c->dat || !c->genmap || !c->chrstat || !c->geno) {
        fclose(f);
        epi_container_free(c);
        return EPIH_ERR_IO;
    }
    int ok = fread(c->dat, sizeof(ped), (size_t)c->nid, f) == (size_t)c->nid &&
             fread(c->genmap, sizeof(map), (size_t)c->nsnp, f) == (size_t)c->nsnp &&
             fread(c->chrstat, sizeof(chromosome), (size_t)c->nchr, f) == (size_t)c->nchr &&
             fread(c->geno, sizeof(int32_t), npack1 * (size_t)c->nsnp, f) == npack1 * (size_t)c->nsnp;
    fclose(f);
    /* Name fields are printed with %s: each must end within its 30 bytes. */
    for (int i = 0; ok && i < c->nid; i++)
        ok = memchr(c->dat[i].family, 0, 30) && memchr(c->dat[i].id, 0, 30) && memchr(c->dat[i].paternal, 0, 30) &&
             memchr(c->dat[i].maternal, 0, 30);
    for (int i = 0; ok && i < c->nsnp; i++) ok = memchr(c->genmap[i].chr, 0, 30) && memchr(c->genmap[i].snpname, 0, 30);
    for (int i = 0; ok && i < c->nchr; i++)
        ok = memchr(c->chrstat[i].chrname, 0, 30) && c->chrstat[i].chrsize >= 0 && c->chrstat[i].chrstart >= 0 &&
             (long long)c->chrstat[i].chrstart + c->chrstat[i].chrsize <= c->nsnp;
    if (!ok) {
        epi_container_free(c);
        return EPIH_ERR_FORMAT;
    }
    return EPIH_OK;
}

int epi_write_container(const char *path, const epi_container *c)
{
    if (c->nid <= 0 || c->nsnp <= 0 || c->nchr <= 0) return EPIH_ERR_ARG;
    FILE *f = fopen(path, "wb");
    if (!f) return EPIH_ERR_IO;
    int npack = c->nid / 16, remain = c->nid % 16;
    int hdr[5] = {c->nid, c->nsnp, npack, remain, c->nchr};
    size_t npack1 = (size_t)npack + (remain > 0);
    int ok = fwrite(SIG, 1, 6, f) == 6 && fwrite(hdr, sizeof(int), 5, f) == 5 &&
             fwrite(c->dat, sizeof(ped), (size_t)c->nid, f) == (size_t)c->nid &&
             fwrite(c->genmap, sizeof(map), (size_t)c->nsnp, f) == (size_t)c->nsnp &&
             fwrite(c->chrstat, sizeof(chromosome), (size_t)c->nchr, f) == (size_t)c->nchr &&
             fwrite(c->geno, sizeof(int32_t), npack1 * (size_t)c->nsnp, f) == npack1 * (size_t)c->nsnp;
    if (fclose(f) != 0) ok = 0;
    return ok ? EPIH_OK : EPIH_ERR_IO;
}
