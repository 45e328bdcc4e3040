/*
 * Host scan driver: mirror of the reference's perform_scan
 * (src/EpiOpenCL.c:1312-1686) over the episcan.h C ABI.
 *
 * Same flow: phenotype from ped records (:1364-1368), resume from the output
 * file or write the header (:1440-1471), then every chr x chr block in the
 * reference order, each block appended to the output file as
 * "# i x j : n interactions" followed by its rows (:1616-1660). Differences:
 *   - one device scan per block replaces the chunk-by-chunk launches (:1576-1603);
 *   - rows of a block are sorted by (SNP1, SNP2) instead of slot order;
 *   - a resumed scan continues with every block after the last complete one
 *     (the reference never resets chr2 for later chromosomes, :1540);
 *   - a block whose rows were cut short by an interrupted append is removed
 *     from the file and scanned again;
 *   - errors are returned, never exit().
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "episcan.h"
#include "epihost.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void epi_run_opts_default(epi_run_opts *o)
{
    memset(o, 0, sizeof(*o));
    o->test = 'f';
    o->t1 = 6.5f;
    o->t2 = 10.5f;
    o->device = 0;
    o->form = EPI_FORM_MFMA;
    o->perm = 0;
    o->datafile = "";
    o->phenfile = "";
    o->quiet = 0;
}

const char *epih_strerror(int code)
{
    switch (code) {
    case EPIH_OK: return "ok";
    case EPIH_ERR_IO: return "file I/O error";
    case EPIH_ERR_FORMAT: return "malformed file";
    case EPIH_ERR_MISMATCH: return "output file belongs to a different analysis";
    case EPIH_ERR_ARG: return "invalid argument";
    case EPIH_ERR_DONE: return "scan already completed";
    default: return code < 0 && code > -100 ? epi_strerror(code) : "unknown error";
    }
}

int epi_format_row(char *buf, size_t n, int test, int snp1, const char *name1, const char *chr1, int snp2,
                   const char *name2, const char *chr2, int df1, int df2, float F, float F2)
{
    if (test == 'f')
        return snprintf(buf, n, "%d\t%s\t%s\t%d\t%s\t%s\t%d\t%d\t%f\n", snp1 + 1, name1, chr1, snp2 + 1, name2, chr2,
                        df1, df2, F);
    return snprintf(buf, n, "%d\t%s\t%s\t%d\t%s\t%s\t%d\t%d\t%f\t%f\n", snp1 + 1, name1, chr1, snp2 + 1, name2, chr2,
                    df1, df2, F, F2);
}

/* Header fields written by a fresh scan (src/EpiOpenCL.c:1459-1469). */
static int write_header(const char *filename, const epi_run_opts *o, int nid, int nsnp)
{
    FILE *f = fopen(filename, "wt");
    if (!f) return EPIH_ERR_IO;
    fprintf(f, "Datafile: %s %s\n", o->datafile ? o->datafile : "", o->phenfile ? o->phenfile : "");
    fprintf(f, "Number of individuals: %d\nNumber of SNPs: %d\nPermutation: %d\nThreshold: %f %f\n", nid, nsnp,
            o->perm, o->t1, o->t2);
    if (o->test == 'f') {
        fprintf(f, "Test: full\n\n");
        fprintf(f, "SNP1\tSNP1name\tChr1\tSNP2\tSNP2name\tChr2\tdf1\tdf2\tFval\n");
    } else {
        fprintf(f, "Test: interactions + full\n\n");
        fprintf(f, "SNP1\tSNP1name\tChr1\tSNP2\tSNP2name\tChr2\tdf1\tdf2\tFval\tFint\n");
    }
    return fclose(f) == 0 ? EPIH_OK : EPIH_ERR_IO;
}

int epi_resume(const char *filename, const char *datafile, int nid, int nsnp, int perm, int test, float t1,
               float t2, epi_resume_state *st)
{
    memset(st, 0, sizeof(*st));
    FILE *f = fopen(filename, "rt");
    if (!f) return EPIH_OK; /* fresh analysis */
    char line[4096], tok[4096];
    int v, c1 = -1, c2 = -1, k;
    float th1, th2;
    char x;
    /* header (src/EpiOpenCL.c:1196-1245) */
    if (!fgets(line, sizeof line, f)) { fclose(f); return EPIH_OK; } /* empty file: fresh */
    tok[0] = 0;
    if (sscanf(line, "Datafile: %4095s", tok) < 1) tok[0] = 0;
    if (strcmp(tok, datafile ? datafile : "") != 0) { fclose(f); return EPIH_ERR_MISMATCH; }
    if (!fgets(line, sizeof line, f) || sscanf(line, "Number of individuals: %d", &v) != 1 || v != nid) { fclose(f); return EPIH_ERR_MISMATCH; }
    if (!fgets(line, sizeof line, f) || sscanf(line, "Number of SNPs: %d", &v) != 1 || v != nsnp) { fclose(f); return EPIH_ERR_MISMATCH; }
    if (!fgets(line, sizeof line, f) || sscanf(line, "Permutation: %d", &v) != 1 || v != perm) { fclose(f); return EPIH_ERR_MISMATCH; }
    if (!fgets(line, sizeof line, f) || sscanf(line, "Threshold: %f %f", &th1, &th2) != 2) { fclose(f); return EPIH_ERR_FORMAT; }
    if (th1 != t1) fprintf(stderr, "Warning: Resuming analysis with different full threshold, changed from %f to %f\n", th1, t1);
    if (th2 != t2) fprintf(stderr, "Warning: Resuming analysis with different interaction threshold, changed from %f to %f\n", th2, t2);
    if (!fgets(line, sizeof line, f) || sscanf(line, "Test: %c", &x) != 1) { fclose(f); return EPIH_ERR_FORMAT; }
    if ((x == 'f' ? 'f' : 'i') != test) { fclose(f); return EPIH_ERR_MISMATCH; }
    if (!fgets(line, sizeof line, f) || !fgets(line, sizeof line, f)) {
        fclose(f);
        st->exists = 1;
        return EPIH_OK; /* header only */
    }
    st->exists = 1;
    long good_end = ftell(f);
    int64_t hits = 0;
    int blocks = 0;
    for (;;) {
        long hdr_at = ftell(f);
        if (!fgets(line, sizeof line, f)) break;
        int a, b;
        if (sscanf(line, "# %d x %d : %d interactions", &a, &b, &k) != 3) break;
        int rows = 0;
        while (rows < k && fgets(line, sizeof line, f)) {
            size_t L = strlen(line);
            if (L == 0 || line[L - 1] != '\n') break; /* cut-off row */
            rows++;
        }
        if (rows < k) {
            good_end = hdr_at;
            break;
        }
        hits += k;
        blocks++;
        c1 = a;
        c2 = b;
        good_end = ftell(f);
    }
    fclose(f);
    /* drop a partial trailing block */
    if (truncate(filename, good_end) != 0) return EPIH_ERR_IO;
    st->tothits = hits;
    st->blocks = blocks;
    if (blocks == 0) {
        st->chr1 = 0;
        st->chr2 = 0;
    } else if (c1 == c2) {
        st->chr1 = c1 + 1;
        st->chr2 = 0;
    } else {
        st->chr1 = c1;
        st->chr2 = c2 + 1;
    }
    return EPIH_OK;
}

int epi_perform_scan(const epi_run_opts *opt, const int32_t *geno, const chromosome *chrstat, const ped *dat,
                     const map *genmap, int nid, int nsnp, int npack, int remain, int nchr, const char *filename,
                     epi_run_report *rep)
{
    double t0 = now_s();
    epi_run_report local;
    if (!rep) rep = &local;
    memset(rep, 0, sizeof(*rep));
    if (!opt || !geno || !chrstat || !dat || !genmap || !filename || nid < 2 || nsnp < 2 || nchr < 1)
        return EPIH_ERR_ARG;
    if (npack != nid / 16 || remain != nid % 16) return EPIH_ERR_ARG;
    if (opt->test != 'f' && opt->test != 'i') return EPIH_ERR_ARG;

    /* phenotype (src/EpiOpenCL.c:1364-1368) */
    float *phen = (float *)malloc(sizeof(float) * (size_t)nid);
    int32_t *cs = (int32_t *)malloc(sizeof(int32_t) * (size_t)nchr);
    int32_t *cz = (int32_t *)malloc(sizeof(int32_t) * (size_t)nchr);
    epi_ctx *ctx = NULL;
    epi_hit *hits = NULL;
    FILE *out = NULL;
    int rc = EPIH_OK;
    if (!phen || !cs || !cz) { rc = EPIH_ERR_IO; goto done; }
    for (int i = 0; i < nid; i++) phen[i] = dat[i].phen;
    for (int i = 0; i < nchr; i++) {
        cs[i] = chrstat[i].chrstart;
        cz[i] = chrstat[i].chrsize;
    }

    /* resume or fresh header (src/EpiOpenCL.c:1440-1471) */
    epi_resume_state st;
    rc = epi_resume(filename, opt->datafile, nid, nsnp, opt->perm, opt->test, opt->t1, opt->t2, &st);
    if (rc) goto done;
    int chrint = nchr * (nchr + 1) / 2;
    int ccount = st.chr1 * (st.chr1 + 1) / 2 + st.chr2;
    rep->total_hits = st.tothits;
    rep->first_block = ccount;
    if (st.exists && st.blocks > 0) {
        if (ccount == chrint) {
            if (!opt->quiet) printf("Scan already completed!\n\n");
            rc = EPIH_OK;
            goto done;
        }
        if (ccount > chrint) { rc = EPIH_ERR_MISMATCH; goto done; }
        if (!opt->quiet) printf("Resuming from %d/%d, %lld hits found so far...\n\n", ccount + 1, chrint, (long long)st.tothits);
    } else if (!st.exists) {
        if ((rc = write_header(filename, opt, nid, nsnp))) goto done;
        ccount = 0;
    }

    /* device (src/EpiOpenCL.c:1474-1524) */
    if ((rc = epi_open(opt->device, &ctx))) goto done;
    if ((rc = epi_set_genotypes(ctx, geno, nid, nsnp)) ||
        (rc = epi_set_phenotypes_ex(ctx, phen, 1, nid, opt->safe ? EPI_PHEN_NINE_IS_VALUE : 0)) ||
        (rc = epi_set_chromosomes(ctx, cs, cz, nchr))) {
        if (!opt->quiet) fprintf(stderr, "episcan: %s\n", epi_last_error(ctx));
        goto done;
    }
    if (!opt->quiet) {
        char name[256];
        epi_device_name(ctx, name, sizeof name);
        printf("Using device %d: %s\n\n", opt->device, name);
    }

    epi_scan_params p;
    epi_scan_params_default(&p);
    p.test = opt->test;
    p.form = opt->form;
    p.t1 = opt->t1;
    p.t2 = opt->t2;
    p.safe = opt->safe;
    int64_t hcap = 0;
    for (int i = 0, blk = 0; i < nchr; i++) {
        for (int j = 0; j <= i; j++, blk++) {
            if (blk < ccount) continue;
            p.blk_begin = blk;
            p.blk_end = blk + 1;
            if (!opt->quiet) {
                printf("%d/%d\tScanning %s vs %s\n", blk + 1, chrint, chrstat[i].chrname, chrstat[j].chrname);
                fflush(stdout);
            }
            if ((rc = epi_scan(ctx, &p))) {
                if (!opt->quiet) fprintf(stderr, "episcan: %s\n", epi_last_error(ctx));
                goto done;
            }
            int64_t nh = 0, npairs = 0;
            int32_t ovf = 0, nl = 0;
            double ms = 0.0;
            epi_hit_count(ctx, &nh, &ovf);
            epi_last_scan_stats(ctx, &ms, &npairs, &nl);
            rep->kernel_s += ms * 1e-3;
            rep->pairs += npairs;
            if (ovf) {
                /* grow the device buffer and rescan this block */
                if ((rc = epi_set_hit_capacity(ctx, nh + 1024)) || (rc = epi_scan(ctx, &p))) goto done;
                epi_hit_count(ctx, &nh, &ovf);
                if (ovf) { rep->overflow = 1; rc = EPI_ERR_NOMEM; goto done; }
            }
            if (nh > hcap) {
                free(hits);
                hcap = nh;
                hits = (epi_hit *)malloc(sizeof(epi_hit) * (size_t)(hcap > 0 ? hcap : 1));
                if (!hits) { rc = EPIH_ERR_IO; goto done; }
            }
            int64_t got = 0;
            if ((rc = epi_get_hits(ctx, hits, NULL, NULL, NULL, nh, &got))) goto done;
            out = fopen(filename, "a");
            if (!out) { rc = EPIH_ERR_IO; goto done; }
            fprintf(out, "# %d x %d : %d interactions\n", i, j, (int)got);
            char row[512];
            for (int64_t t = 0; t < got; t++) {
                const epi_hit *h = &hits[t];
                epi_format_row(row, sizeof row, opt->test, h->snp1, genmap[h->snp1].snpname, chrstat[i].chrname,
                               h->snp2, genmap[h->snp2].snpname, chrstat[j].chrname, h->df1, h->df2, h->F, h->F2);
                fputs(row, out);
            }
            if (fclose(out) != 0) { out = NULL; rc = EPIH_ERR_IO; goto done; }
            out = NULL;
            rep->new_hits += got;
            rep->total_hits += got;
            rep->blocks_done++;
            if (!opt->quiet) {
                printf("\n\t%d interactions found in %.3f s (device)\n\n", (int)got, ms * 1e-3);
                fflush(stdout);
            }
        }
    }
    if (!opt->quiet)
        printf("\nProgramme finished in %.1f s\n%lld epistatic interactions discovered\n", now_s() - t0,
               (long long)rep->total_hits);
done:
    if (out) fclose(out);
    if (ctx) epi_close(ctx);
    free(hits);
    free(phen);
    free(cs);
    free(cz);
    rep->wall_s = now_s() - t0;
    return rc;
}

void epi_permute_order(unsigned seed, int nid, int32_t *order)
{
    for (int i = 0; i < nid; i++) order[i] = i;
    srand(seed);
    for (int i = nid; i > 1; i--) {
        int r = rand() % i;
        int32_t t = order[r];
        order[r] = order[i - 1];
        order[i - 1] = t;
    }
}

void epi_permute_phen(unsigned seed, int nid, const ped *dat, float *phen_out)
{
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nid > 0 ? nid : 1));
    if (!order) {
        for (int i = 0; i < nid; i++) phen_out[i] = dat[i].phen;
        return;
    }
    epi_permute_order(seed, nid, order);
    for (int i = 0; i < nid; i++) phen_out[i] = dat[order[i]].phen;
    free(order);
}

int epi_write_perm_file(const char *path, int nid, const ped *dat, const int32_t *order, const float *phen)
{
    FILE *f = fopen(path, "w");
    if (!f) return EPIH_ERR_IO;
    fprintf(f, "Family ID Phenotype\n");
    for (int i = 0; i < nid; i++) {
        const int k = order[i];
        fprintf(f, "%s %s %f\n", dat[k].family, dat[k].id, phen ? phen[k] : dat[k].phen);
    }
    return fclose(f) == 0 ? EPIH_OK : EPIH_ERR_IO;
}

/* Output file name of permutation k from a format with one %d. */
static int perm_name(char *buf, size_t n, const char *fmt, int k)
{
    const char *pc = strstr(fmt, "%d");
    if (!pc || strstr(pc + 2, "%") || strchr(fmt, '%') != pc) return EPIH_ERR_ARG;
    int m = snprintf(buf, n, "%.*s%d%s", (int)(pc - fmt), fmt, k, pc + 2);
    return m > 0 && (size_t)m < n ? EPIH_OK : EPIH_ERR_ARG;
}

int epi_perform_scan_multi(const epi_run_opts *opt, const int32_t *geno, const chromosome *chrstat,
                           const map *genmap, int nid, int nsnp, int npack, int remain, int nchr, int nscan,
                           const float *phen, const int *perms, const char *const *phenfiles,
                           const char *const *outfiles, epi_run_report *rep)
{
    double t0 = now_s();
    epi_run_report local;
    if (!rep) rep = &local;
    memset(rep, 0, sizeof(*rep));
    if (!opt || !geno || !chrstat || !genmap || !phen || !outfiles || nid < 2 || nsnp < 2 || nchr < 1 || nscan < 1)
        return EPIH_ERR_ARG;
    if (npack != nid / 16 || remain != nid % 16) return EPIH_ERR_ARG;
    if (opt->test != 'f' && opt->test != 'i') return EPIH_ERR_ARG;
    for (int s = 0; s < nscan; s++) {
        if (!outfiles[s]) return EPIH_ERR_ARG;
        for (int t = 0; t < s; t++)
            if (strcmp(outfiles[s], outfiles[t]) == 0) return EPIH_ERR_ARG;
        FILE *f = fopen(outfiles[s], "r");
        if (f) {
            fclose(f);
            return EPIH_ERR_DONE;
        }
    }
    int32_t *cs = (int32_t *)malloc(sizeof(int32_t) * (size_t)nchr);
    int32_t *cz = (int32_t *)malloc(sizeof(int32_t) * (size_t)nchr);
    epi_ctx *ctx = NULL;
    epi_batch *b = NULL;
    epi_hit *hits = NULL;
    int32_t *blk = NULL;
    FILE *out = NULL;
    int rc = EPIH_OK;
    if (!cs || !cz) { rc = EPIH_ERR_IO; goto done; }
    for (int i = 0; i < nchr; i++) {
        cs[i] = chrstat[i].chrstart;
        cz[i] = chrstat[i].chrsize;
    }
    if ((rc = epi_open(opt->device, &ctx)) || (rc = epi_batch_create(ctx, &b))) goto done;
    int32_t d = 0, s0 = 0;
    if ((rc = epi_batch_add_dataset(b, geno, nid, nsnp, cs, cz, nchr, &d)) ||
        (rc = epi_batch_add_phenotypes_ex(b, d, phen, nscan, opt->safe ? EPI_PHEN_NINE_IS_VALUE : 0, &s0))) {
        if (!opt->quiet) fprintf(stderr, "episcan: %s\n", epi_last_error(ctx));
        goto done;
    }
    epi_scan_params p;
    epi_scan_params_default(&p);
    p.test = opt->test;
    p.form = opt->form;
    p.t1 = opt->t1;
    p.t2 = opt->t2;
    p.safe = opt->safe;
    if (!opt->quiet) {
        printf("Scanning %d phenotypes in one batched launch\n", nscan);
        fflush(stdout);
    }
    if ((rc = epi_batch_scan(b, &p))) {
        if (!opt->quiet) fprintf(stderr, "episcan: %s\n", epi_last_error(ctx));
        goto done;
    }
    int64_t tot = 0;
    int32_t ovf = 0;
    epi_batch_hit_count(b, -1, &tot, &ovf);
    if (ovf) {
        /* grow the shared device buffer and rescan */
        if ((rc = epi_batch_set_hit_capacity(b, tot + 1024)) || (rc = epi_batch_scan(b, &p))) goto done;
        epi_batch_hit_count(b, -1, &tot, &ovf);
        if (ovf) { rep->overflow = 1; rc = EPI_ERR_NOMEM; goto done; }
    }
    {
        double ms = 0.0;
        int64_t pairs = 0, items = 0;
        int32_t caps = 0, grid = 0;
        epi_batch_stats(b, &ms, &pairs, &items, &caps, &grid);
        rep->kernel_s = ms * 1e-3;
        rep->pairs = pairs;
    }
    const int nblk = nchr * (nchr + 1) / 2;
    for (int sc = 0; sc < nscan; sc++) {
        const int s = s0 + sc;
        int64_t nh = 0, got = 0;
        if ((rc = epi_batch_hit_count(b, s, &nh, NULL))) goto done;
        free(hits);
        free(blk);
        hits = (epi_hit *)malloc(sizeof(epi_hit) * (size_t)(nh > 0 ? nh : 1));
        blk = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nh > 0 ? nh : 1));
        if (!hits || !blk) { rc = EPIH_ERR_IO; goto done; }
        if ((rc = epi_batch_get_hits(b, s, hits, NULL, NULL, blk, nh, &got))) goto done;
        epi_run_opts o = *opt;
        o.perm = perms ? perms[sc] : 0;
        o.phenfile = phenfiles && phenfiles[sc] ? phenfiles[sc] : "";
        if ((rc = write_header(outfiles[sc], &o, nid, nsnp))) goto done;
        out = fopen(outfiles[sc], "a");
        if (!out) { rc = EPIH_ERR_IO; goto done; }
        int64_t t = 0;
        for (int i = 0, bi = 0; i < nchr; i++)
            for (int j = 0; j <= i; j++, bi++) {
                int64_t t1 = t;
                while (t1 < got && blk[t1] == bi) t1++;
                fprintf(out, "# %d x %d : %d interactions\n", i, j, (int)(t1 - t));
                char row[512];
                for (; t < t1; t++) {
                    const epi_hit *h = &hits[t];
                    epi_format_row(row, sizeof row, opt->test, h->snp1, genmap[h->snp1].snpname, chrstat[i].chrname,
                                   h->snp2, genmap[h->snp2].snpname, chrstat[j].chrname, h->df1, h->df2, h->F,
                                   h->F2);
                    fputs(row, out);
                }
            }
        if (fclose(out) != 0) { out = NULL; rc = EPIH_ERR_IO; goto done; }
        out = NULL;
        rep->new_hits += got;
        rep->total_hits += got;
        rep->blocks_done += nblk;
        if (!opt->quiet) printf("%lld interactions -> %s\n", (long long)got, outfiles[sc]);
    }
    if (!opt->quiet)
        printf("\nProgramme finished in %.1f s (device %.3f s)\n", now_s() - t0, rep->kernel_s);
done:
    if (out) fclose(out);
    if (b) epi_batch_destroy(b);
    if (ctx) epi_close(ctx);
    free(hits);
    free(blk);
    free(cs);
    free(cz);
    rep->wall_s = now_s() - t0;
    return rc;
}

int epi_perform_scan_perms(const epi_run_opts *opt, const int32_t *geno, const chromosome *chrstat, const ped *dat,
                           const map *genmap, int nid, int nsnp, int npack, int remain, int nchr, int first, int last,
                           const char *fmt, epi_run_report *rep)
{
    epi_run_report local;
    if (!rep) rep = &local;
    memset(rep, 0, sizeof(*rep));
    if (!opt || !dat || !fmt || nid < 2 || first < 1 || last < first) return EPIH_ERR_ARG;
    const int np = last - first + 1;
    float *phen = (float *)malloc(sizeof(float) * (size_t)nid * (size_t)np);
    int *perms = (int *)malloc(sizeof(int) * (size_t)np);
    char **names = (char **)calloc((size_t)np, sizeof(char *));
    const char **pf = (const char **)calloc((size_t)np, sizeof(char *));
    int rc = EPIH_OK;
    if (!phen || !perms || !names || !pf) { rc = EPIH_ERR_IO; goto done; }
    for (int k = first; k <= last; k++) {
        const int s = k - first;
        names[s] = (char *)malloc(4096);
        if (!names[s]) { rc = EPIH_ERR_IO; goto done; }
        if ((rc = perm_name(names[s], 4096, fmt, k))) goto done;
        perms[s] = k;
        pf[s] = opt->phenfile;
        epi_permute_phen((unsigned)k, nid, dat, phen + (size_t)s * nid);
    }
    rc = epi_perform_scan_multi(opt, geno, chrstat, genmap, nid, nsnp, npack, remain, nchr, np, phen, perms, pf,
                                (const char *const *)names, rep);
done:
    if (names)
        for (int s = 0; s < np; s++) free(names[s]);
    free(names);
    free(pf);
    free(perms);
    free(phen);
    return rc;
}
