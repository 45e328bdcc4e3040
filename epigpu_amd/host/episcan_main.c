/*
 * episcan - analysis-mode CLI (mirror of src/main.c:58-188, :283-331).
 *
 *   episcan -A <data.epigpu> <output.txt> [options]
 *     -t f|i     8df test, or 8df + 4df interaction test (default f)
 *     -F x       F threshold of the 8df test (default 6.5)
 *     -I x       F threshold of the 4df test (default 10.5)
 *     -f p       set -F from a p-value: F^-1(1-p; 8, N-9) (extension)
 *     -g p       set -I from a p-value: F^-1(1-p; 4, N-9) (extension)
 *     -d n       HIP device (default 0)
 *     -p n       permutation n: shuffle phenotypes with srand(n) and write perm<n>.txt,
 *                as the reference (src/main.c:309-318, :336-357)
 *     -p a-b     permutations a..b in one batched device launch, one output file
 *                <output.txt>.perm<k> and one perm<k>.txt per permutation (extension)
 *     -P file    phenotype file, one value per individual (src/main.c:172-181, :191-222).
 *                Repeatable (extension): several files scan in one batched device
 *                launch, file j (1-based) into <output.txt>.P<j>; with -p, every file's
 *                permutations too, into <output.txt>.P<j>.perm<k> and perm<k>.P<j>.txt
 *     -m plain|mfma  kernel form (default mfma) (extension)
 *     -s 0|1     safe mode: -9 is a phenotype value and test i drops nfac > 6, as
 *                the reference's safe kernels (src/EpiOpenCL.c:505-963)
 *     -i n       accepted for compatibility and ignored (tiles replace launch chunks)
 * Options come in pairs, as in the reference (src/main.c:63).
 *
 * Data mode (src/readplink.c:1296-1435):
 *   episcan -Dr[c][m][n] <file.ped> <file.map> <out.epigpu>   PLINK text to a container
 *   episcan -Db[c][m][n] <prefix> <out.epigpu>                prefix.bed/.bim/.fam to a container
 *   episcan -De <file.ped> <file.map> <in.epigpu>             container back to PLINK text
 *     c  clean: drop low call-rate / low-frequency SNPs, then low call-rate individuals
 *        and individuals with phenotype -9
 *     m  impute missing calls from a random individual's call
 *     n  replace the phenotype with N(0,1) draws
 * The random steps use the stream the reference's unseeded rand() gives (seed 1).
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "episcan.h"
#include "epihost.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: episcan -A <data.epigpu> <output.txt> [-t f|i] [-F x] [-I x] [-f p] [-g p]\n"
            "               [-d n] [-p n | -p a-b] [-P file ...] [-m plain|mfma] [-s 0|1]\n"
            "       episcan -Dr[c][m][n] <file.ped> <file.map> <out.epigpu>\n"
            "       episcan -Db[c][m][n] <prefix> <out.epigpu>\n"
            "       episcan -De <file.ped> <file.map> <in.epigpu>\n");
}

static int ingest_fail(int rc)
{
    const char *m = epi_plink_error();
    fprintf(stderr, "episcan: %s%s%s\n", epih_strerror(rc), m[0] ? ": " : "", m);
    return 1;
}

static int datamode(int argc, char **argv)
{
    int txt = 0, bed = 0, extract = 0, flags = 0;
    for (const char *q = argv[1] + 2; *q; q++) {
        switch (*q) {
        case 'r': txt = 1; break;
        case 'b': bed = 1; break;
        case 'e': extract = 1; break;
        case 'c': flags |= EPI_PLINK_CLEAN; break;
        case 'm': flags |= EPI_PLINK_IMPUTE; break;
        case 'n': flags |= EPI_PLINK_RNORM; break;
        default:
            fprintf(stderr, "data mode flag '%c' not supported (r, b, e, c, m, n)\n", *q);
            return 1;
        }
    }
    if (txt + bed + extract != 1 || (extract && flags) || argc != (bed ? 4 : 5)) {
        usage();
        return 1;
    }
    int rc;
    if (extract) {
        epi_container c;
        if ((rc = epi_read_container(argv[4], &c))) {
            fprintf(stderr, "%s: %s\n", argv[4], epih_strerror(rc));
            return 1;
        }
        rc = epi_write_ped_map(&c, argv[2], argv[3]);
        epi_container_free(&c);
        if (rc) return ingest_fail(rc);
        printf("Extracted %s to %s and %s\n", argv[4], argv[2], argv[3]);
        return 0;
    }
    epi_plink d;
    if (bed) {
        const char *pre = argv[2];
        size_t n = strlen(pre) + 5;
        char *fb = (char *)malloc(n), *fm = (char *)malloc(n), *ff = (char *)malloc(n);
        if (!fb || !fm || !ff) return 1;
        snprintf(fb, n, "%s.bed", pre);
        snprintf(fm, n, "%s.bim", pre);
        snprintf(ff, n, "%s.fam", pre);
        rc = epi_read_bed(fb, fm, ff, &d);
        free(fb);
        free(fm);
        free(ff);
    } else {
        rc = epi_read_ped_map(argv[2], argv[3], &d);
    }
    if (rc) return ingest_fail(rc);
    printf("%d individuals, %d SNPs, %d chromosomes\n", d.nid, d.nsnp, d.nchr);
    epi_plink_report rep;
    if ((rc = epi_plink_process(&d, flags, 1u, &rep))) {
        epi_plink_free(&d);
        return ingest_fail(rc);
    }
    if (flags & EPI_PLINK_CLEAN)
        printf("Cleaning kept %d of %d SNPs and %d of %d individuals\n", rep.snps_kept, rep.snps_in, rep.ids_kept,
               rep.ids_in);
    if (flags & EPI_PLINK_IMPUTE) printf("%lld missing calls imputed\n", (long long)rep.imputed);
    if (flags & EPI_PLINK_RNORM) printf("Phenotype replaced by N(0,1) draws\n");
    epi_container c;
    rc = epi_plink_to_container(&d, &c);
    epi_plink_free(&d);
    if (rc) return ingest_fail(rc);
    const char *out = argv[argc - 1];
    rc = epi_write_container(out, &c);
    epi_container_free(&c);
    if (rc) {
        fprintf(stderr, "%s: %s\n", out, epih_strerror(rc));
        return 1;
    }
    printf("Stored as %s\n", out);
    return 0;
}

static int all_digits(const char *s)
{
    if (!*s) return 0;
    for (; *s; s++)
        if (!isdigit((unsigned char)*s)) return 0;
    return 1;
}

static const char *base_name(const char *p)
{
    const char *b = p;
    for (; *p; p++)
        if (*p == '/' || *p == '\\') b = p + 1;
    return b;
}

/* 1/2-coded phenotype check (src/readplink.c:118-147). */
static int is_binary(int nid, const ped *dat)
{
    int c1 = 0, c2 = 0, other = 0;
    for (int i = 0; i < nid; i++) {
        if (dat[i].phen == 1.0f) c1++;
        else if (dat[i].phen == 2.0f) c2++;
        else other++;
    }
    return other == 0 && c1 + c2 == nid;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && strncmp(argv[1], "-D", 2) == 0) return datamode(argc, argv);
    if (argc < 4 || strcmp(argv[1], "-A") != 0 || (argc & 1) != 0) {
        usage();
        return 1;
    }
    const char *bin = argv[2], *outfile = argv[3];
    epi_run_opts o;
    epi_run_opts_default(&o);
    enum { MAXP = 256 };
    const char *phenfiles[MAXP];
    int nphenf = 0;
    double pf = -1.0, pi = -1.0;
    int pfirst = 0, plast = 0;
    for (int i = 4; i < argc; i += 2) {
        const char *f = argv[i], *v = argv[i + 1];
        if (f[0] != '-' || !f[1] || f[2]) { usage(); return 1; }
        switch (f[1]) {
        case 't':
            if (v[0] != 'f' && v[0] != 'i') { fprintf(stderr, "Test selection: f or i\n"); return 1; }
            o.test = v[0];
            break;
        case 'F': o.t1 = (float)atof(v); if (o.t1 == 0.0f) { fprintf(stderr, "numeric -F required\n"); return 1; } break;
        case 'I': o.t2 = (float)atof(v); if (o.t2 == 0.0f) { fprintf(stderr, "numeric -I required\n"); return 1; } break;
        case 'f': pf = atof(v); break;
        case 'g': pi = atof(v); break;
        case 'd': if (!all_digits(v)) { fprintf(stderr, "numeric -d required\n"); return 1; } o.device = atoi(v); break;
        case 'p': {
            const char *dash = strchr(v, '-');
            if (dash) {
                char a[32];
                size_t la = (size_t)(dash - v);
                if (la == 0 || la >= sizeof a) { fprintf(stderr, "-p a-b: numeric a and b required\n"); return 1; }
                memcpy(a, v, la);
                a[la] = 0;
                if (!all_digits(a) || !all_digits(dash + 1)) { fprintf(stderr, "-p a-b: numeric a and b required\n"); return 1; }
                pfirst = atoi(a);
                plast = atoi(dash + 1);
                if (pfirst < 1 || plast < pfirst) { fprintf(stderr, "-p a-b needs 1 <= a <= b\n"); return 1; }
            } else {
                if (!all_digits(v)) { fprintf(stderr, "numeric -p required\n"); return 1; }
                o.perm = atoi(v);
            }
            break;
        }
        case 'P':
            if (nphenf == MAXP) { fprintf(stderr, "at most %d phenotype files\n", MAXP); return 1; }
            phenfiles[nphenf++] = v;
            break;
        case 'm':
            if (strcmp(v, "plain") == 0) o.form = EPI_FORM_PLAIN;
            else if (strcmp(v, "mfma") == 0) o.form = EPI_FORM_MFMA;
            else { fprintf(stderr, "-m plain|mfma\n"); return 1; }
            break;
        case 'i': {
            int k = atoi(v);
            if (!all_digits(v) || k < 16 || k % 16) { fprintf(stderr, "Iteration size must be a multiple of 16\n"); return 1; }
            break;
        }
        case 's': if (strcmp(v, "0") && strcmp(v, "1")) { fprintf(stderr, "-s 0|1\n"); return 1; } o.safe = v[0] == '1'; break;
        default: usage(); return 1;
        }
    }
    epi_container c;
    int rc = epi_read_container(bin, &c);
    if (rc) {
        fprintf(stderr, "%s: %s\n", bin, epih_strerror(rc));
        return 1;
    }
    /* phenotype sources: the container's phenotypes, or one vector per -P file */
    const int nsrc = nphenf ? nphenf : 1;
    float *src = (float *)malloc(sizeof(float) * (size_t)c.nid * (size_t)nsrc);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)c.nid);
    if (!src || !order) { fprintf(stderr, "out of memory\n"); return 1; }
    for (int i = 0; i < c.nid; i++) src[i] = c.dat[i].phen;
    for (int jf = 0; jf < nphenf; jf++) {
        FILE *pfh = fopen(phenfiles[jf], "rt");
        if (!pfh) { fprintf(stderr, "Separate phenotype file does not exist: %s\n", phenfiles[jf]); return 1; }
        for (int i = 0; i < c.nid; i++) {
            float t;
            if (fscanf(pfh, "%f", &t) != 1) { fprintf(stderr, "phenotype file too short: %s\n", phenfiles[jf]); return 1; }
            src[(size_t)jf * c.nid + i] = t;
        }
        fclose(pfh);
    }
    for (int jf = 0; jf < nsrc; jf++) {
        for (int i = 0; i < c.nid; i++) c.dat[i].phen = src[(size_t)jf * c.nid + i];
        if (is_binary(c.nid, c.dat)) {
            fprintf(stderr, "Error: Binary phenotype!\nepiGPU can only use quantitative phenotypes for analysis.\n");
            return 1;
        }
    }
    if (pf > 0.0) o.t1 = (float)epi_f_isf(pf, 8.0, (double)c.nid - 9.0);
    if (pi > 0.0) o.t2 = (float)epi_f_isf(pi, 4.0, (double)c.nid - 9.0);
    o.datafile = base_name(bin);
    o.phenfile = nphenf ? phenfiles[0] : "";
    epi_run_report rep;
    if (nsrc == 1 && pfirst == 0) {
        /* one scan, resumable: the reference's path, with its permutation (src/main.c:309-318) */
        for (int i = 0; i < c.nid; i++) c.dat[i].phen = src[i];
        if (o.perm > 0) {
            char pname[64];
            epi_permute_order((unsigned)o.perm, c.nid, order);
            snprintf(pname, sizeof pname, "perm%d.txt", o.perm);
            if (epi_write_perm_file(pname, c.nid, c.dat, order, src)) {
                fprintf(stderr, "cannot write %s\n", pname);
                return 1;
            }
            for (int i = 0; i < c.nid; i++) c.dat[i].phen = src[order[i]];
            printf("Permutation %d\n", o.perm);
            for (int i = 0; i < 10 && i < c.nid; i++) printf("%f\n", c.dat[i].phen);
        }
        rc = epi_perform_scan(&o, c.geno, c.chrstat, c.dat, c.genmap, c.nid, c.nsnp, c.npack, c.remain, c.nchr,
                              outfile, &rep);
    } else {
        /* every (source, permutation) pair is one scan of a single batched launch */
        if (pfirst == 0 && o.perm > 0) pfirst = plast = o.perm;
        const int nperm = pfirst ? plast - pfirst + 1 : 1, nscan = nsrc * nperm;
        float *ph = (float *)malloc(sizeof(float) * (size_t)c.nid * (size_t)nscan);
        int *perms = (int *)malloc(sizeof(int) * (size_t)nscan);
        const char **pfs = (const char **)malloc(sizeof(char *) * (size_t)nscan);
        char **names = (char **)calloc((size_t)nscan, sizeof(char *));
        const size_t nlen = strlen(outfile) + 64;
        if (!ph || !perms || !pfs || !names) { fprintf(stderr, "out of memory\n"); return 1; }
        for (int jf = 0; jf < nsrc; jf++)
            for (int kk = 0; kk < nperm; kk++) {
                const int s = jf * nperm + kk, k = pfirst ? pfirst + kk : 0;
                const float *y = src + (size_t)jf * c.nid;
                float *dst = ph + (size_t)s * c.nid;
                char tag[32] = "", pname[96];
                if (nsrc > 1) snprintf(tag, sizeof tag, ".P%d", jf + 1);
                if (k) {
                    epi_permute_order((unsigned)k, c.nid, order);
                    for (int i = 0; i < c.nid; i++) dst[i] = y[order[i]];
                    snprintf(pname, sizeof pname, "perm%d%s.txt", k, tag);
                    if (epi_write_perm_file(pname, c.nid, c.dat, order, y)) {
                        fprintf(stderr, "cannot write %s\n", pname);
                        return 1;
                    }
                } else {
                    memcpy(dst, y, sizeof(float) * (size_t)c.nid);
                }
                names[s] = (char *)malloc(nlen);
                if (!names[s]) { fprintf(stderr, "out of memory\n"); return 1; }
                if (k) snprintf(names[s], nlen, "%s%s.perm%d", outfile, tag, k);
                else snprintf(names[s], nlen, "%s%s", outfile, tag);
                perms[s] = k;
                pfs[s] = nphenf ? phenfiles[jf] : "";
            }
        rc = epi_perform_scan_multi(&o, c.geno, c.chrstat, c.genmap, c.nid, c.nsnp, c.npack, c.remain, c.nchr, nscan,
                                    ph, perms, pfs, (const char *const *)names, &rep);
        for (int s = 0; s < nscan; s++) free(names[s]);
        free(names);
        free(pfs);
        free(perms);
        free(ph);
    }
    free(order);
    free(src);
    epi_container_free(&c);
    if (rc) {
        fprintf(stderr, "episcan: %s (%d)\n", epih_strerror(rc), rc);
        return 1;
    }
    printf("pairs %lld  device %.3f s  wall %.3f s  hits %lld\n", (long long)rep.pairs, rep.kernel_s, rep.wall_s,
           (long long)rep.total_hits);
    return 0;
}
