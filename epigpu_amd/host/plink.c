/*
 * PLINK ingest: the reference's data mode (src/readplink.c:1296-1435) as
 * library calls.
 *
 *   epi_read_ped_map   text .ped/.map          dims :8-60, readped :234-350, readmap :352-415
 *   epi_read_bed       PLINK 1 .bed/.bim/.fam  planned in README.md:12-13
 *   epi_plink_process  cleangeno :417-497, guessmissing :499-520, rnorm :533-552
 *   epi_plink_to_container                     pack :1066-1102
 *   epi_write_ped_map  extractfrombinary :150-224
 *
 * Parsing. The .ped is read twice: one pass counts rows and the first row's
 * columns, the second parses rows in blocks of 64 individuals and scatters
 * each block into the SNP-major code matrix (64 contiguous bytes per SNP).
 * While parsing, a call is stored as 1 (heterozygote), 3 (first allele '0')
 * or, for a homozygote, as an allele value >= 4 (the allele character for a
 * .ped, 0x80/0x81 for A1/A2 of a .bed). One pass per SNP then turns allele
 * values into classes: the first homozygote in individual order becomes 0,
 * the other one 2. SNPs are independent there, so that pass runs in parallel.
 *
 * Randomness. The imputation and phenotype steps draw from glibc's
 * random_r() on a private TYPE_3 state seeded like srand(seed), i.e. the same
 * sequence rand() gives after srand(seed), without touching the process-wide
 * rand() state.
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "epihost.h"

#define NAME_MAX_LEN 29   /* ped/map/chromosome name fields are char[30] */
#define BLOCK_ROWS 64
#define BED_A1 0x80
#define BED_A2 0x81

static __thread char g_msg[256];

const char *epi_plink_error(void) { return g_msg; }

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    return code;
}

void epi_plink_free(epi_plink *d)
{
    if (!d) return;
    free(d->dat);
    free(d->genmap);
    free(d->chrstat);
    free(d->codes);
    free(d->snpchr);
    memset(d, 0, sizeof(*d));
}

/* ------------------------------------------------------------- tokens -- */

static int is_sep(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

/* Next whitespace-separated token of [*p, end). */
static int next_tok(const char **p, const char *end, const char **tok, size_t *len)
{
    const char *s = *p;
    while (s < end && is_sep(*s)) s++;
    if (s >= end) return 0;
    const char *e = s;
    while (e < end && !is_sep(*e)) e++;
    *tok = s;
    *len = (size_t)(e - s);
    *p = e;
    return 1;
}

static size_t count_toks(const char *s, const char *end)
{
    size_t n = 0;
    const char *t;
    size_t l;
    while (next_tok(&s, end, &t, &l)) n++;
    return n;
}

static int blank(const char *s, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!is_sep(s[i])) return 0;
    return 1;
}

static int copy_name(char *dst, const char *tok, size_t len, const char *what, long line)
{
    if (len > NAME_MAX_LEN)
        return fail(EPIH_ERR_FORMAT, "line %ld: %s '%.*s' longer than %d bytes", line, what, (int)len, tok,
                    NAME_MAX_LEN);
    memcpy(dst, tok, len);
    dst[len] = 0;
    return EPIH_OK;
}

/* Whole token as a number: strtol base 0 (what %li accepts), strtol base 10
 * (%d), strtod (%lf) and strtof (%f). */
static int tok_long(const char *tok, size_t len, int base, long *v)
{
    char buf[64];
    if (len == 0 || len >= sizeof buf) return 0;
    memcpy(buf, tok, len);
    buf[len] = 0;
    char *e;
    errno = 0;
    *v = strtol(buf, &e, base);
    return *e == 0 && errno == 0;
}

static int tok_double(const char *tok, size_t len, double *v)
{
    char buf[64];
    if (len == 0 || len >= sizeof buf) return 0;
    memcpy(buf, tok, len);
    buf[len] = 0;
    char *e;
    *v = strtod(buf, &e);
    return *e == 0;
}

static int tok_float(const char *tok, size_t len, float *v)
{
    char buf[64];
    if (len == 0 || len >= sizeof buf) return 0;
    memcpy(buf, tok, len);
    buf[len] = 0;
    char *e;
    *v = strtof(buf, &e);
    return *e == 0;
}

/* The six leading columns of a .ped row or a .fam line. */
static int parse_person(const char **p, const char *end, ped *r, long line)
{
    const char *t;
    size_t l;
    char *fields[4] = {r->family, r->id, r->paternal, r->maternal};
    static const char *names[4] = {"family id", "individual id", "paternal id", "maternal id"};
    for (int k = 0; k < 4; k++) {
        if (!next_tok(p, end, &t, &l)) return fail(EPIH_ERR_FORMAT, "line %ld: fewer than 6 columns", line);
        int rc = copy_name(fields[k], t, l, names[k], line);
        if (rc) return rc;
    }
    long sex;
    if (!next_tok(p, end, &t, &l) || !tok_long(t, l, 10, &sex))
        return fail(EPIH_ERR_FORMAT, "line %ld: sex is not an integer", line);
    r->sex = (char)sex;
    if (!next_tok(p, end, &t, &l) || !tok_float(t, l, &r->phen))
        return fail(EPIH_ERR_FORMAT, "line %ld: phenotype is not a number", line);
    return EPIH_OK;
}

/* Line reader that skips blank lines. */
typedef struct lines {
    FILE *f;
    char *buf;
    size_t cap;
    long no;
} lines;

static ssize_t next_line(lines *L)
{
    for (;;) {
        ssize_t n = getline(&L->buf, &L->cap, L->f);
        if (n < 0) return -1;
        L->no++;
        if (!blank(L->buf, (size_t)n)) return n;
    }
}

/* --------------------------------------------------------- chromosomes -- */

/* Chromosome list from the per-SNP names (readmap, src/readplink.c:371-409):
 * order of first appearance, chrid = index, chrstart cumulative. */
static int build_chromosomes(epi_plink *d)
{
    d->snpchr = (int32_t *)malloc(sizeof(int32_t) * (size_t)d->nsnp);
    d->chrstat = (chromosome *)calloc((size_t)d->nsnp, sizeof(chromosome));
    if (!d->snpchr || !d->chrstat) return fail(EPIH_ERR_IO, "out of memory");
    int n = 0;
    for (int s = 0; s < d->nsnp; s++) {
        const char *name = d->genmap[s].chr;
        if (n > 0 && strcmp(name, d->chrstat[n - 1].chrname) == 0) {
            d->chrstat[n - 1].chrsize++;
            d->snpchr[s] = n - 1;
            continue;
        }
        for (int c = 0; c < n - 1; c++)
            if (strcmp(name, d->chrstat[c].chrname) == 0)
                return fail(EPIH_ERR_FORMAT, "SNP %d: chromosome %s appears again after other chromosomes", s + 1,
                            name);
        chromosome *c = &d->chrstat[n];
        c->chrid = n;
        strcpy(c->chrname, name);
        c->chrsize = 1;
        c->chrstart = s;
        d->snpchr[s] = n++;
    }
    d->nchr = n;
    return EPIH_OK;
}

/* ------------------------------------------------------ class recoding -- */

/* Allele values of one SNP (1 het, 3 missing, >= 4 homozygous allele) to
 * classes, and the SNP's allele names (readped, src/readplink.c:292-340).
 * het0/het1: the last heterozygote's alleles (0 if none). bed_names: names of
 * BED_A1/BED_A2, or NULL when the value is the allele character itself. */
static int recode_snp(uint8_t *g, int nid, char het0, char het1, const char *bed_names, map *m, int s)
{
    int p1 = 0;
    while (p1 < nid && g[p1] < 4) p1++;
    m->allelename[0] = het0;
    m->allelename[1] = het1;
    if (p1 == nid) return EPIH_OK;  /* only heterozygous or missing calls */
    const uint8_t v1 = g[p1];
    const char c1 = bed_names ? bed_names[v1 - BED_A1] : (char)v1;
    if (het0 == 0) {
        m->allelename[0] = c1;
    } else if (c1 == het1) {
        m->allelename[0] = c1;
        m->allelename[1] = het0;
    }
    int v2 = -1;
    for (int p = p1; p < nid; p++) {
        const uint8_t v = g[p];
        if (v < 4) continue;
        if (v == v1) {
            g[p] = 0;
        } else if (v2 < 0 || v == v2) {
            v2 = v;
            g[p] = 2;
        } else {
            return fail(EPIH_ERR_FORMAT, "SNP %d: more than two homozygous alleles", s + 1);
        }
    }
    return EPIH_OK;
}

static int recode_all(epi_plink *d, const char *het, const char *bed_names)
{
    int rc = EPIH_OK;
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < d->nsnp; s++) {
        const char *bn = bed_names ? bed_names + 2 * (size_t)s : NULL;
        int r = recode_snp(d->codes + (size_t)s * d->nid, d->nid, het[2 * (size_t)s], het[2 * (size_t)s + 1], bn,
                           &d->genmap[s], s);
        if (r) {
#pragma omp critical
            rc = r;
        }
    }
    return rc;
}

/* ------------------------------------------------------------- .ped/.map -- */

static int read_map(const char *mapfile, epi_plink *d)
{
    lines L = {fopen(mapfile, "r"), NULL, 0, 0};
    if (!L.f) return fail(EPIH_ERR_IO, "cannot open %s", mapfile);
    int rc = EPIH_OK, s = 0;
    ssize_t n;
    while ((n = next_line(&L)) >= 0) {
        if (s >= d->nsnp) { rc = fail(EPIH_ERR_FORMAT, "%s: more than %d SNPs", mapfile, d->nsnp); break; }
        const char *p = L.buf, *end = L.buf + n, *t;
        size_t l;
        map *m = &d->genmap[s];
        if (count_toks(p, end) != 4) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: need 4 columns", mapfile, L.no); break; }
        next_tok(&p, end, &t, &l);
        if ((rc = copy_name(m->chr, t, l, "chromosome", L.no))) break;
        next_tok(&p, end, &t, &l);
        if ((rc = copy_name(m->snpname, t, l, "SNP name", L.no))) break;
        next_tok(&p, end, &t, &l);
        if (!tok_double(t, l, &m->gd)) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: bad genetic distance", mapfile, L.no); break; }
        next_tok(&p, end, &t, &l);
        if (!tok_long(t, l, 0, &m->pd)) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: bad position", mapfile, L.no); break; }
        s++;
    }
    if (!rc && s != d->nsnp) rc = fail(EPIH_ERR_FORMAT, "%s has %d SNPs, the .ped %d", mapfile, s, d->nsnp);
    free(L.buf);
    fclose(L.f);
    return rc;
}

/* Scatter rows [i0, i0+nr) of a row-major block into the SNP-major codes. */
static void scatter_block(epi_plink *d, const uint8_t *blk, int i0, int nr)
{
#pragma omp parallel for schedule(static) if (d->nsnp > 4096)
    for (int s = 0; s < d->nsnp; s++) {
        uint8_t *dst = d->codes + (size_t)s * d->nid + i0;
        for (int r = 0; r < nr; r++) dst[r] = blk[(size_t)r * d->nsnp + s];
    }
}

int epi_read_ped_map(const char *pedfile, const char *mapfile, epi_plink *d)
{
    g_msg[0] = 0;
    if (!pedfile || !mapfile || !d) return fail(EPIH_ERR_ARG, "null argument");
    memset(d, 0, sizeof(*d));
    /* pass 1: rows and columns (dims) */
    lines L = {fopen(pedfile, "r"), NULL, 0, 0};
    if (!L.f) return fail(EPIH_ERR_IO, "cannot open %s", pedfile);
    ssize_t n;
    size_t cols = 0;
    long rows = 0;
    while ((n = next_line(&L)) >= 0) {
        if (rows == 0) cols = count_toks(L.buf, L.buf + n);
        rows++;
    }
    int rc = EPIH_OK;
    if (rows < 2 || cols < 8 || (cols - 6) % 2 || rows > 0x7fffffff || (cols - 6) / 2 > 0x7fffffff) {
        rc = fail(EPIH_ERR_FORMAT, "%s: %ld rows of %zu columns; need 6 + 2 per SNP", pedfile, rows, cols);
        goto out;
    }
    d->nid = (int)rows;
    d->nsnp = (int)((cols - 6) / 2);
    d->dat = (ped *)calloc((size_t)d->nid, sizeof(ped));
    d->genmap = (map *)calloc((size_t)d->nsnp, sizeof(map));
    d->codes = (uint8_t *)malloc((size_t)d->nsnp * (size_t)d->nid);
    char *het = (char *)calloc((size_t)d->nsnp, 2);
    uint8_t *blk = (uint8_t *)malloc((size_t)BLOCK_ROWS * (size_t)d->nsnp);
    if (!d->dat || !d->genmap || !d->codes || !het || !blk) {
        rc = fail(EPIH_ERR_IO, "out of memory");
        goto out2;
    }
    /* pass 2: parse rows (readped) */
    rewind(L.f);
    L.no = 0;
    int i = 0, nr = 0;
    while ((n = next_line(&L)) >= 0) {
        const char *p = L.buf, *end = L.buf + n, *t;
        size_t l;
        if (i >= d->nid) {  /* the file grew between the two passes */
            rc = fail(EPIH_ERR_FORMAT, "%s changed while it was read", pedfile);
            goto out2;
        }
        if ((rc = parse_person(&p, end, &d->dat[i], L.no))) goto out2;
        uint8_t *row = blk + (size_t)nr * d->nsnp;
        for (int s = 0; s < d->nsnp; s++) {
            const char *ta, *tb;
            size_t la, lb;
            if (!next_tok(&p, end, &ta, &la) || !next_tok(&p, end, &tb, &lb)) {
                rc = fail(EPIH_ERR_FORMAT, "%s line %ld: %d SNPs, the first row %d", pedfile, L.no, s, d->nsnp);
                goto out2;
            }
            if (la != 1 || lb != 1) {
                rc = fail(EPIH_ERR_FORMAT, "%s line %ld: SNP %d allele is not one character", pedfile, L.no, s + 1);
                goto out2;
            }
            const char a = ta[0], b = tb[0];
            if (a == '0') {
                row[s] = EPI_MISS;
            } else if (a != b) {
                het[2 * (size_t)s] = a;
                het[2 * (size_t)s + 1] = b;
                row[s] = 1;
            } else {
                row[s] = (uint8_t)a;
            }
        }
        if (next_tok(&p, end, &t, &l)) {
            rc = fail(EPIH_ERR_FORMAT, "%s line %ld: more columns than the first row", pedfile, L.no);
            goto out2;
        }
        i++;
        if (++nr == BLOCK_ROWS) {
            scatter_block(d, blk, i - nr, nr);
            nr = 0;
        }
    }
    if (nr) scatter_block(d, blk, i - nr, nr);
    if (i != d->nid) {
        rc = fail(EPIH_ERR_FORMAT, "%s changed while it was read", pedfile);
        goto out2;
    }
    if ((rc = recode_all(d, het, NULL))) goto out2;
    if ((rc = read_map(mapfile, d))) goto out2;
    rc = build_chromosomes(d);
out2:
    free(het);
    free(blk);
out:
    free(L.buf);
    fclose(L.f);
    if (rc) epi_plink_free(d);
    return rc;
}

/* ------------------------------------------------------- .bed/.bim/.fam -- */

int epi_read_bed(const char *bedfile, const char *bimfile, const char *famfile, epi_plink *d)
{
    g_msg[0] = 0;
    if (!bedfile || !bimfile || !famfile || !d) return fail(EPIH_ERR_ARG, "null argument");
    memset(d, 0, sizeof(*d));
    int rc = EPIH_OK;
    char *het = NULL, *names = NULL;
    uint8_t *raw = NULL;
    FILE *bf = NULL;
    /* .fam: the six person columns */
    lines L = {fopen(famfile, "r"), NULL, 0, 0};
    if (!L.f) return fail(EPIH_ERR_IO, "cannot open %s", famfile);
    ssize_t n;
    int cap = 0;
    while ((n = next_line(&L)) >= 0) {
        if (d->nid == cap) {
            cap = cap ? 2 * cap : 1024;
            ped *nd = (ped *)realloc(d->dat, sizeof(ped) * (size_t)cap);
            if (!nd) { rc = fail(EPIH_ERR_IO, "out of memory"); goto out; }
            d->dat = nd;
        }
        ped *r = &d->dat[d->nid];
        memset(r, 0, sizeof(*r));
        const char *p = L.buf, *end = L.buf + n;
        if (count_toks(p, end) != 6) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: need 6 columns", famfile, L.no); goto out; }
        if ((rc = parse_person(&p, end, r, L.no))) goto out;
        d->nid++;
    }
    fclose(L.f);
    L.f = NULL;
    if (d->nid < 2) { rc = fail(EPIH_ERR_FORMAT, "%s: fewer than 2 individuals", famfile); goto out; }
    /* .bim: chr, name, cM, bp, A1, A2 */
    L.f = fopen(bimfile, "r");
    L.no = 0;
    if (!L.f) { rc = fail(EPIH_ERR_IO, "cannot open %s", bimfile); goto out; }
    cap = 0;
    while ((n = next_line(&L)) >= 0) {
        if (d->nsnp == cap) {
            cap = cap ? 2 * cap : 4096;
            map *nm = (map *)realloc(d->genmap, sizeof(map) * (size_t)cap);
            char *nn = (char *)realloc(names, 2 * (size_t)cap);
            if (nm) d->genmap = nm;
            if (nn) names = nn;
            if (!nm || !nn) { rc = fail(EPIH_ERR_IO, "out of memory"); goto out; }
        }
        map *m = &d->genmap[d->nsnp];
        memset(m, 0, sizeof(*m));
        const char *p = L.buf, *end = L.buf + n, *t;
        size_t l;
        if (count_toks(p, end) != 6) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: need 6 columns", bimfile, L.no); goto out; }
        next_tok(&p, end, &t, &l);
        if ((rc = copy_name(m->chr, t, l, "chromosome", L.no))) goto out;
        next_tok(&p, end, &t, &l);
        if ((rc = copy_name(m->snpname, t, l, "SNP name", L.no))) goto out;
        next_tok(&p, end, &t, &l);
        if (!tok_double(t, l, &m->gd)) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: bad genetic distance", bimfile, L.no); goto out; }
        next_tok(&p, end, &t, &l);
        if (!tok_long(t, l, 0, &m->pd)) { rc = fail(EPIH_ERR_FORMAT, "%s line %ld: bad position", bimfile, L.no); goto out; }
        for (int k = 0; k < 2; k++) {
            next_tok(&p, end, &t, &l);
            /* an allele spelled "0" is PLINK's missing allele: its calls read back as missing */
            names[2 * (size_t)d->nsnp + k] = (l == 1 && t[0] == '0') ? 0 : t[0];
        }
        d->nsnp++;
    }
    if (d->nsnp < 1) { rc = fail(EPIH_ERR_FORMAT, "%s: no SNPs", bimfile); goto out; }
    /* .bed: magic 6c 1b, mode 01 (SNP-major), then ceil(nid/4) bytes per SNP,
     * individual k in bits [2(k%4)+1 : 2(k%4)] of byte k/4:
     * 00 A1/A1, 01 missing, 10 A1/A2, 11 A2/A2 */
    bf = fopen(bedfile, "rb");
    if (!bf) { rc = fail(EPIH_ERR_IO, "cannot open %s", bedfile); goto out; }
    const size_t bps = ((size_t)d->nid + 3) / 4;
    unsigned char magic[3];
    if (fread(magic, 1, 3, bf) != 3 || magic[0] != 0x6c || magic[1] != 0x1b) {
        rc = fail(EPIH_ERR_FORMAT, "%s is not a PLINK .bed file", bedfile);
        goto out;
    }
    if (magic[2] != 1) { rc = fail(EPIH_ERR_FORMAT, "%s is individual-major; only SNP-major .bed is read", bedfile); goto out; }
    raw = (uint8_t *)malloc(bps * (size_t)d->nsnp);
    d->codes = (uint8_t *)malloc((size_t)d->nid * (size_t)d->nsnp);
    het = (char *)calloc((size_t)d->nsnp, 2);
    if (!raw || !d->codes || !het) { rc = fail(EPIH_ERR_IO, "out of memory"); goto out; }
    if (fread(raw, 1, bps * (size_t)d->nsnp, bf) != bps * (size_t)d->nsnp || fgetc(bf) != EOF) {
        rc = fail(EPIH_ERR_FORMAT, "%s size does not match %d individuals x %d SNPs", bedfile, d->nid, d->nsnp);
        goto out;
    }
#pragma omp parallel for schedule(static)
    for (int s = 0; s < d->nsnp; s++) {
        const uint8_t *b = raw + bps * (size_t)s;
        uint8_t *g = d->codes + (size_t)s * d->nid;
        const char a1 = names[2 * (size_t)s], a2 = names[2 * (size_t)s + 1];
        int any_het = 0;
        for (int k = 0; k < d->nid; k++) {
            const int v = (b[k >> 2] >> (2 * (k & 3))) & 3;
            uint8_t c;
            if (v == 1) c = EPI_MISS;
            else if (v == 0) c = a1 ? BED_A1 : EPI_MISS;
            else if (v == 3) c = a2 ? BED_A2 : EPI_MISS;
            else if (a1 == 0) c = EPI_MISS;  /* written "0 A2": first allele '0' */
            else { c = 1; any_het = 1; }
            g[k] = c;
        }
        if (any_het) {
            het[2 * (size_t)s] = a1;
            het[2 * (size_t)s + 1] = a2 ? a2 : '0';
        }
    }
    if ((rc = recode_all(d, het, names))) goto out;
    rc = build_chromosomes(d);
out:
    if (L.f) fclose(L.f);
    if (bf) fclose(bf);
    free(L.buf);
    free(raw);
    free(het);
    free(names);
    if (rc) epi_plink_free(d);
    return rc;
}

/* --------------------------------------------------------- processing -- */

typedef struct rng {
    struct random_data rd;
    char state[128];
} rng;

static void rng_seed(rng *r, unsigned seed)
{
    memset(r, 0, sizeof(*r));
    initstate_r(seed, r->state, sizeof r->state, &r->rd);
}

static int rng_next(rng *r)
{
    int32_t v;
    random_r(&r->rd, &v);
    return v;
}

/* cleangeno (src/readplink.c:417-497), with its float arithmetic. */
static int clean(epi_plink *d, epi_plink_report *rep)
{
    const int nid = d->nid, nsnp = d->nsnp;
    int *keep_s = (int *)malloc(sizeof(int) * (size_t)nsnp);
    int *keep_i = (int *)malloc(sizeof(int) * (size_t)nid);
    if (!keep_s || !keep_i) { free(keep_s); free(keep_i); return fail(EPIH_ERR_IO, "out of memory"); }
    int ns = 0;
    for (int s = 0; s < nsnp; s++) {
        const uint8_t *g = d->codes + (size_t)s * nid;
        int k = 0, sum = 0;
        for (int i = 0; i < nid; i++)
            if (g[i] != EPI_MISS) {
                sum += g[i];
                k++;
            }
        const float callrate = (float)k / (float)nid;
        const float freq = (float)sum / (float)(2 * k);
        if (freq > 0.01 && callrate > 0.9 && d->genmap[s].pd > 0) keep_s[ns++] = s;
    }
    int ni = 0;
    for (int i = 0; i < nid; i++) {
        int k = 0;
        for (int j = 0; j < ns; j++) k += d->codes[(size_t)keep_s[j] * nid + i] != EPI_MISS;
        const float callrate = (float)k / (float)ns;
        if (callrate > 0.9 && d->dat[i].phen != -9) keep_i[ni++] = i;
    }
    size_t o = 0;
    for (int j = 0; j < ns; j++) {
        const uint8_t *g = d->codes + (size_t)keep_s[j] * nid;  /* keep_s[j] >= j: reads stay ahead of writes */
        for (int t = 0; t < ni; t++) d->codes[o++] = g[keep_i[t]];
        d->genmap[j] = d->genmap[keep_s[j]];
        d->snpchr[j] = d->snpchr[keep_s[j]];
    }
    for (int t = 0; t < ni; t++) d->dat[t] = d->dat[keep_i[t]];
    /* chromosome list keeps every entry; sizes and starts over the kept SNPs */
    for (int c = 0; c < d->nchr; c++) d->chrstat[c].chrsize = 0;
    for (int j = 0; j < ns; j++) d->chrstat[d->snpchr[j]].chrsize++;
    for (int c = 0; c < d->nchr; c++) d->chrstat[c].chrstart = c ? d->chrstat[c - 1].chrstart + d->chrstat[c - 1].chrsize : 0;
    d->nid = ni;
    d->nsnp = ns;
    rep->snps_kept = ns;
    rep->ids_kept = ni;
    free(keep_s);
    free(keep_i);
    if (ns < 1 || ni < 2) return fail(EPIH_ERR_FORMAT, "cleaning left %d SNPs and %d individuals", ns, ni);
    return EPIH_OK;
}

/* guessmissing (src/readplink.c:499-520). */
static int impute(epi_plink *d, rng *r, epi_plink_report *rep)
{
    const int nid = d->nid;
    for (int s = 0; s < d->nsnp; s++) {
        const uint8_t *g = d->codes + (size_t)s * nid;
        int any = 0;
        for (int i = 0; i < nid && !any; i++) any = g[i] != EPI_MISS;
        if (!any) return fail(EPIH_ERR_FORMAT, "SNP %d has no calls to impute from", s + 1);
    }
    for (int s = 0; s < d->nsnp; s++) {
        uint8_t *g = d->codes + (size_t)s * nid;
        for (int i = 0; i < nid; i++) {
            if (g[i] != EPI_MISS) continue;
            rep->imputed++;
            while (g[i] == EPI_MISS) g[i] = g[rng_next(r) % nid];
        }
    }
    return EPIH_OK;
}

/* rnorm (src/readplink.c:533-552): Box-Muller from two uniforms, with the
 * reference's float arithmetic and its constant 3.14159. */
static void rnorm(epi_plink *d, rng *r)
{
    for (int i = 0; i < d->nid; i++) {
        const float u1 = (float)rng_next(r) / RAND_MAX;
        const float u2 = (float)rng_next(r) / RAND_MAX;
        const float radius = sqrtf(-2 * logf(u1));
        const float c = cosf(2 * 3.14159 * u2);
        d->dat[i].phen = radius * c;
    }
}

int epi_plink_process(epi_plink *d, int flags, unsigned seed, epi_plink_report *rep)
{
    g_msg[0] = 0;
    epi_plink_report local;
    if (!rep) rep = &local;
    memset(rep, 0, sizeof(*rep));
    if (!d || !d->codes || (flags & ~(EPI_PLINK_CLEAN | EPI_PLINK_IMPUTE | EPI_PLINK_RNORM)))
        return fail(EPIH_ERR_ARG, "bad arguments");
    rep->snps_in = rep->snps_kept = d->nsnp;
    rep->ids_in = rep->ids_kept = d->nid;
    int rc;
    if ((flags & EPI_PLINK_CLEAN) && (rc = clean(d, rep))) return rc;
    rng r;
    rng_seed(&r, seed);
    if ((flags & EPI_PLINK_IMPUTE) && (rc = impute(d, &r, rep))) return rc;
    if (flags & EPI_PLINK_RNORM) rnorm(d, &r);
    return EPIH_OK;
}

int epi_plink_to_container(const epi_plink *d, epi_container *c)
{
    g_msg[0] = 0;
    if (!d || !c || !d->codes || d->nid < 1 || d->nsnp < 1 || d->nchr < 1) return fail(EPIH_ERR_ARG, "empty dataset");
    memset(c, 0, sizeof(*c));
    int npack1;
    epi_layout(d->nid, &c->npack, &c->remain, &npack1);
    c->nid = d->nid;
    c->nsnp = d->nsnp;
    c->nchr = d->nchr;
    c->dat = (ped *)malloc(sizeof(ped) * (size_t)d->nid);
    c->genmap = (map *)malloc(sizeof(map) * (size_t)d->nsnp);
    c->chrstat = (chromosome *)malloc(sizeof(chromosome) * (size_t)d->nchr);
    c->geno = (int32_t *)malloc(sizeof(int32_t) * (size_t)npack1 * (size_t)d->nsnp);
    if (!c->dat || !c->genmap || !c->chrstat || !c->geno) {
        epi_container_free(c);
        return fail(EPIH_ERR_IO, "out of memory");
    }
    memcpy(c->dat, d->dat, sizeof(ped) * (size_t)d->nid);
    memcpy(c->genmap, d->genmap, sizeof(map) * (size_t)d->nsnp);
    memcpy(c->chrstat, d->chrstat, sizeof(chromosome) * (size_t)d->nchr);
    epi_pack((const int8_t *)d->codes, d->nid, d->nsnp, c->geno);
    return EPIH_OK;
}

/* ------------------------------------------------------------ extract -- */

int epi_write_ped_map(const epi_container *c, const char *pedfile, const char *mapfile)
{
    g_msg[0] = 0;
    if (!c || !pedfile || !mapfile || !c->geno || c->nid < 1 || c->nsnp < 1) return fail(EPIH_ERR_ARG, "bad arguments");
    FILE *mf = fopen(mapfile, "w");
    if (!mf) return fail(EPIH_ERR_IO, "cannot create %s", mapfile);
    for (int s = 0; s < c->nsnp; s++) {
        const map *m = &c->genmap[s];
        fprintf(mf, "%s %s %f %ld\n", m->chr, m->snpname, m->gd, m->pd);
    }
    if (fclose(mf) != 0) return fail(EPIH_ERR_IO, "cannot write %s", mapfile);
    /* 1/2-coded phenotype (isbinary, src/readplink.c:118-147) */
    int binary = 1;
    for (int i = 0; i < c->nid && binary; i++) binary = c->dat[i].phen == 1.0f || c->dat[i].phen == 2.0f;
    int8_t *codes = (int8_t *)malloc((size_t)c->nid * (size_t)c->nsnp);
    char *row = (char *)malloc(4 * (size_t)c->nsnp + 4 * 32 + 256);
    FILE *pf = fopen(pedfile, "w");
    int rc = EPIH_OK;
    if (!codes || !row) rc = fail(EPIH_ERR_IO, "out of memory");
    else if (!pf) rc = fail(EPIH_ERR_IO, "cannot create %s", pedfile);
    if (!rc) {
        epi_unpack(c->geno, c->nid, c->nsnp, codes);
        for (int i = 0; i < c->nid && !rc; i++) {
            const ped *r = &c->dat[i];
            if (binary)
                fprintf(pf, "%s %s %s %s %d %d", r->family, r->id, r->paternal, r->maternal, (int)r->sex, (int)r->phen);
            else
                fprintf(pf, "%s %s %s %s %d %f", r->family, r->id, r->paternal, r->maternal, (int)r->sex, r->phen);
            size_t o = 0;
            for (int s = 0; s < c->nsnp; s++) {
                const char *an = c->genmap[s].allelename;
                char x, y;
                switch (codes[(size_t)s * c->nid + i]) {
                case 0: x = an[0]; y = an[0]; break;
                case 1: x = an[0]; y = an[1]; break;
                case 2: x = an[1]; y = an[1]; break;
                default: x = '0'; y = '0'; break;
                }
                row[o++] = ' ';
                row[o++] = x;
                row[o++] = ' ';
                row[o++] = y;
            }
            row[o++] = '\n';
            if (fwrite(row, 1, o, pf) != o) rc = fail(EPIH_ERR_IO, "cannot write %s", pedfile);
        }
    }
    if (pf && fclose(pf) != 0 && !rc) rc = fail(EPIH_ERR_IO, "cannot write %s", pedfile);
    free(codes);
    free(row);
    return rc;
}
