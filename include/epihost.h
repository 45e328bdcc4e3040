/*
 * epihost.h - C host side of the MI355X exhaustive pairwise epistasis scan.
 *
 * libepihost.so is plain C (gcc). It holds everything of the scan that runs on
 * the host: the 2-bit genotype codec, the epiGPU packed-binary container, the
 * seeded synthetic workload generator, the F-distribution helpers that turn a
 * p-value into an F threshold, the block driver that mirrors the reference's
 * perform_scan (chr x chr blocks, text output, resume), and the CLI. It calls
 * the GPU only through the C ABI declared in episcan.h.
 *
 * Reference interfaces mirrored here (paths relative to the epiGPU tree):
 *   ped / map / chromosome           src/readplink.h:17-39
 *   pack / unpack                    src/readplink.c:1066-1143
 *   createpackedbinary / readpacked  src/readplink.c:1145-1273
 *   perform_scan                     src/EpiOpenCL.c:1312-1686, src/epiopencl.h:21
 *   resume                           src/EpiOpenCL.c:1187-1274
 *   PLINK data mode (ingest, QC)     src/readplink.c:8-560, :1296-1435
 *   analysis-mode defaults           src/main.c:244-249
 */
#ifndef EPIHOST_H
#define EPIHOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Records of the epiGPU container. Field order, sizes and padding follow    */
/* src/readplink.h:17-39 on x86-64 gcc so that a container written by the    */
/* reference reads back byte for byte: sizeof = 128, 88 and 44 bytes.        */
/* ------------------------------------------------------------------------ */
typedef struct ped {
    char family[30];
    char id[30];
    char paternal[30];
    char maternal[30];
    char sex;
    float phen;
} ped;

typedef struct map {
    char chr[30];
    char snpname[30];
    double gd;
    long pd;
    char allelename[2];
} map;

typedef struct chromosome {
    int chrid;
    char chrname[30];
    int chrsize;
    int chrstart;
} chromosome;

/* Genotype code for a missing call (src/readplink.h:9). */
#define EPI_MISS 3

/* ------------------------------------------------------------------------ */
/* Packed genotype codec (layout A1).                                        */
/* 16 genotypes per int32, SNP-major, npack1 = ceil(nid/16) words per SNP.   */
/* Individual 16w+t of a full word sits in bits [31-2t : 30-2t]. The tail    */
/* word holds remain = nid % 16 genotypes right-aligned: individual          */
/* 16*npack+t at bits [2(remain-1-t)+1 : 2(remain-1-t)].                     */
/* ------------------------------------------------------------------------ */
void epi_layout(int nid, int *npack, int *remain, int *npack1);
/* codes: nsnp x nid bytes (SNP-major), values 0,1,2,3. words: nsnp*npack1. */
void epi_pack(const int8_t *codes, int nid, int nsnp, int32_t *words);
void epi_unpack(const int32_t *words, int nid, int nsnp, int8_t *codes);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic workload generator.                                     */
/* Streams: xoshiro256** seeded through splitmix64 from (seed, purpose, id), */
/* so every SNP and every phenotype has its own reproducible stream and the  */
/* generator runs in parallel over SNPs.                                     */
/*   genotypes: per SNP s (purpose 1, id s): maf = maf_lo + (maf_hi-maf_lo)*U;*/
/*     per individual: g = [U<maf] + [U<maf]; then, if miss > 0, a third U   */
/*     marks the call missing (code 3) when U < miss.                        */
/*   phenotype r (purpose 2, id r): one value per individual drawn from the  */
/*     distribution below; then, if miss > 0, a further U < miss -> NaN.     */
/* ------------------------------------------------------------------------ */
enum epi_phen_dist {
    EPI_PHEN_NORMAL = 0,      /* N(0,1), Box-Muller */
    EPI_PHEN_GAMMA2 = 1,      /* Gamma(2,1) = -ln U1 - ln U2 (raw) */
    EPI_PHEN_GAMMA2_STD = 2,  /* (Gamma(2,1) - 2) / sqrt(2) */
    EPI_PHEN_T5 = 3,          /* Student t with 5 df: Z / sqrt(chi2_5 / 5) */
    EPI_PHEN_BERN02 = 4       /* 0/1 with P(1) = 0.2 */
};
void epi_synth_genotypes(uint64_t seed, int nid, int nsnp, double maf_lo,
                         double maf_hi, double miss, int32_t *words);
void epi_synth_phenotype(uint64_t seed, int r, int nid, int dist, double miss,
                         float *phen);
/* Plant npairs interacting pairs (purpose 3): pair (a > b) and a cell
 * (g_a, g_b) drawn uniformly; every individual with that genotype pair and a
 * present phenotype gets +effect. pairs_out: 4 ints per pair (a, b, ga, gb). */
void epi_synth_plant(uint64_t seed, const int32_t *words, int nid, int nsnp,
                     float *phen, int npairs, float effect, int32_t *pairs_out);
/* Fill chromosome records for nchr blocks of the given sizes (names "1".."n"). */
void epi_synth_chromosomes(int nchr, const int *sizes, chromosome *chrstat);

/* ------------------------------------------------------------------------ */
/* F distribution: survival function and its inverse, from the regularised   */
/* incomplete beta function evaluated by its continued fraction (modified    */
/* Lentz). Used to map a p-value to the F threshold the scan takes.          */
/* ------------------------------------------------------------------------ */
double epi_betai(double a, double b, double x);
double epi_f_sf(double F, double d1, double d2);
double epi_f_isf(double p, double d1, double d2);

/* ------------------------------------------------------------------------ */
/* epiGPU packed binary container (src/readplink.c:1145-1273):               */
/* "epiGPU" | int nid,nsnp,npack,remain,nchr | ped[nid] | map[nsnp] |        */
/* chromosome[nchr] | int32 words[npack1*nsnp]. Arrays are malloc'ed; free   */
/* with epi_container_free. Returns 0 or a negative EPIH_ error.             */
/* ------------------------------------------------------------------------ */
typedef struct epi_container {
    int nid, nsnp, npack, remain, nchr;
    ped *dat;
    map *genmap;
    chromosome *chrstat;
    int32_t *geno;
} epi_container;
int epi_read_container(const char *path, epi_container *c);
int epi_write_container(const char *path, const epi_container *c);
void epi_container_free(epi_container *c);

/* Host-side error codes (negative). */
#define EPIH_OK 0
#define EPIH_ERR_IO -101
#define EPIH_ERR_FORMAT -102
#define EPIH_ERR_MISMATCH -103
#define EPIH_ERR_ARG -104
#define EPIH_ERR_DONE -105
const char *epih_strerror(int code);

/* ------------------------------------------------------------------------ */
/* PLINK ingest: the reference's data mode (src/readplink.c:1296-1435).      */
/*                                                                          */
/* Genotype classes follow readped (src/readplink.c:234-350): per SNP the    */
/* allele of the first homozygous call in individual order is class 0, the   */
/* other homozygote class 2, a heterozygote 1 and a call whose first allele  */
/* is '0' missing (3). map.allelename holds the alleles as readped leaves    */
/* them: the last heterozygote's pair in file order, reordered so that the   */
/* class-0 allele comes first; without heterozygotes {class-0 allele, 0}.    */
/* Chromosomes are numbered in order of first appearance and must occupy     */
/* contiguous SNP runs (src/readplink.c:352-415).                            */
/* Divergences from the reference, all on input it mishandles: tab-separated */
/* .ped files are read (readped's separator set is " /t"), ragged rows, odd  */
/* allele counts, names over 29 bytes, non-numeric sex or phenotype,          */
/* non-contiguous chromosomes and a third homozygous allele are errors, and  */
/* a SNP without any homozygous call keeps its heterozygote allele names      */
/* (readped reads past the SNP there).                                      */
/* ------------------------------------------------------------------------ */
typedef struct epi_plink {
    int nid, nsnp, nchr;
    ped *dat;              /* nid records */
    map *genmap;           /* nsnp records */
    chromosome *chrstat;   /* nchr records */
    uint8_t *codes;        /* nsnp x nid, SNP-major, classes 0..3 */
    int32_t *snpchr;       /* nsnp chromosome indices */
} epi_plink;

/* Text .ped (6 columns, then 2 single-character alleles per SNP, space or   */
/* tab separated) and .map (chr, name, genetic distance, base-pair position).*/
/* dims/readped/readmap: src/readplink.c:8-60, :234-350, :352-415.           */
int epi_read_ped_map(const char *pedfile, const char *mapfile, epi_plink *d);
/* PLINK 1 binary fileset (SNP-major .bed, .bim, .fam): the README's planned */
/* direct reader (README.md:12-13). Classes and allele names are those       */
/* epi_read_ped_map gives for the .ped/.map PLINK writes for the same data   */
/* (heterozygotes as "A1 A2", missing as "0 0"); alleles longer than one     */
/* character are named by their first character.                            */
int epi_read_bed(const char *bedfile, const char *bimfile, const char *famfile, epi_plink *d);

/* Data-mode steps, applied in the reference's order (src/readplink.c:1346-1369):
 * EPI_PLINK_CLEAN   cleangeno (:417-497): keep SNPs with allele-2 frequency >
 *                   0.01, call rate > 0.9 and position > 0, then individuals
 *                   with call rate > 0.9 over the kept SNPs and phenotype != -9
 * EPI_PLINK_IMPUTE  guessmissing (:499-520): every missing call copies the call
 *                   of a random individual, rand() % nid, drawn until present
 * EPI_PLINK_RNORM   rnorm (:533-552): phenotypes replaced by Box-Muller draws
 * The random steps share one stream equal to srand(seed); rand() (the
 * reference never seeds it in data mode, so seed 1 reproduces its files). */
#define EPI_PLINK_CLEAN 1
#define EPI_PLINK_IMPUTE 2
#define EPI_PLINK_RNORM 4
typedef struct epi_plink_report {
    int snps_in, snps_kept;
    int ids_in, ids_kept;
    int64_t imputed;
} epi_plink_report;
int epi_plink_process(epi_plink *d, int flags, unsigned seed, epi_plink_report *rep);
/* Packs the codes (layout A1) into a container (pack, src/readplink.c:1066-1102). */
int epi_plink_to_container(const epi_plink *d, epi_container *c);
void epi_plink_free(epi_plink *d);
/* Container back to .ped/.map text (extractfrombinary, src/readplink.c:150-224):
 * 1/2-coded phenotypes print as integers, others with %f. */
int epi_write_ped_map(const epi_container *c, const char *pedfile, const char *mapfile);
/* Message of the last ingest error on this thread ("" if none). */
const char *epi_plink_error(void);

/* ------------------------------------------------------------------------ */
/* Scan driver: mirror of perform_scan (src/EpiOpenCL.c:1312-1686).          */
/* ------------------------------------------------------------------------ */
typedef struct epi_run_opts {
    int test;            /* 'f' (8df) or 'i' (8df + 4df); src/main.c:144-171 */
    float t1, t2;        /* F thresholds; defaults 6.5 / 10.5 (src/main.c:248-249) */
    int device;          /* -d */
    int form;            /* 0 plain popcount kernel, 1 MFMA contraction */
    int perm;            /* permutation number written to the header (-p) */
    const char *datafile;/* name written after "Datafile:" */
    const char *phenfile;/* second token of the Datafile line (-P), may be "" */
    int quiet;           /* 1: no progress lines on stdout */
    int safe;            /* 1: the reference's safe mode (-s 1): -9 is a phenotype value and
                            test 'i' drops nfac > 6 (src/EpiOpenCL.c:505-963, :1486-1499) */
} epi_run_opts;
void epi_run_opts_default(epi_run_opts *o);

typedef struct epi_run_report {
    int64_t total_hits;     /* hits in the file, including resumed blocks */
    int64_t new_hits;       /* hits written by this call */
    int blocks_done;        /* blocks scanned by this call */
    int first_block;        /* first block index scanned (resume point) */
    int overflow;           /* 1 if the device hit buffer overflowed */
    double kernel_s;        /* device time of the scan kernels */
    double wall_s;          /* wall time of the whole call */
    int64_t pairs;          /* pairs tested by this call */
} epi_run_report;

/* Runs every chr x chr block not yet present in `filename` (resume), appends
 * "# i x j : n interactions" sections and hit rows in the reference format,
 * rows sorted by (SNP1, SNP2). geno is layout A1. Returns 0 or an error. */
int epi_perform_scan(const epi_run_opts *opt, const int32_t *geno,
                     const chromosome *chrstat, const ped *dat,
                     const map *genmap, int nid, int nsnp, int npack,
                     int remain, int nchr, const char *filename,
                     epi_run_report *rep);

/* Record order of permutation k: the reference's permute() (src/main.c:336-357)
 * shuffles the ped records with srand(k) and rand() % i from i = nid down to
 * 2 (src/main.c:309-313); order[i] receives the original index of the record
 * that lands at position i. */
void epi_permute_order(unsigned seed, int nid, int32_t *order);

/* Phenotypes of permutation k: phen_out[i] = dat[order[i]].phen. */
void epi_permute_phen(unsigned seed, int nid, const ped *dat, float *phen_out);

/* The side file permute() writes, perm<k>.txt (src/main.c:349-356): a
 * "Family ID Phenotype" line, then "family id phen" (phen with %f) of the
 * permuted records in their new order. phen (may be NULL: dat[].phen) holds
 * the phenotypes by original index, e.g. those of a -P file. */
int epi_write_perm_file(const char *path, int nid, const ped *dat, const int32_t *order, const float *phen);

/* Several phenotype vectors of one dataset scanned together in one batched
 * device launch, sharing the genotype bitplanes: the reference runs one
 * process per phenotype file (-P, src/main.c:172-181, :295-298) and per
 * permutation (-p, :309-312). Scan s scans phen[s*nid .. s*nid + nid-1] and
 * writes a complete output file outfiles[s] in the format of
 * epi_perform_scan, whose header names phenfiles[s] (NULL or "": none) and
 * permutation perms[s] (perms NULL: 0). Existing output files and repeated
 * names are refused (EPIH_ERR_DONE, EPIH_ERR_ARG); this mode does not resume.
 * opt->perm and opt->phenfile are ignored. rep sums over all scans. */
int epi_perform_scan_multi(const epi_run_opts *opt, const int32_t *geno,
                           const chromosome *chrstat, const map *genmap, int nid,
                           int nsnp, int npack, int remain, int nchr, int nscan,
                           const float *phen, const int *perms,
                           const char *const *phenfiles, const char *const *outfiles,
                           epi_run_report *rep);

/* Permutations first..last (1 <= first <= last) of dat[].phen scanned
 * together (epi_perform_scan_multi): the reference runs one process per
 * permutation (-p k, src/main.c:101-112, :309-331). Output file of
 * permutation k is fmt with its one %d replaced by k, with "Permutation: k"
 * and opt->phenfile in its header. opt->perm is ignored. */
int epi_perform_scan_perms(const epi_run_opts *opt, const int32_t *geno,
                           const chromosome *chrstat, const ped *dat,
                           const map *genmap, int nid, int nsnp, int npack,
                           int remain, int nchr, int first, int last,
                           const char *fmt, epi_run_report *rep);

/* Resume state parsed from an existing output file
 * (src/EpiOpenCL.c:1187-1274, with the chr2 restart fixed: a resumed run
 * continues at block (chr1, chr2) and then scans every later block). */
typedef struct epi_resume_state {
    int exists;          /* output file present with a header */
    int chr1, chr2;      /* next block to scan */
    int64_t tothits;     /* hits already in the file */
    int blocks;          /* complete blocks found */
} epi_resume_state;
int epi_resume(const char *filename, const char *datafile, int nid, int nsnp,
               int perm, int test, float t1, float t2, epi_resume_state *st);

/* Reference-format text for one hit row (src/EpiOpenCL.c:1632-1653). */
int epi_format_row(char *buf, size_t n, int test, int snp1, const char *name1,
                   const char *chr1, int snp2, const char *name2,
                   const char *chr2, int df1, int df2, float F, float F2);

#ifdef __cplusplus
}
#endif
#endif /* EPIHOST_H */
