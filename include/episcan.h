/*
 * episcan.h - C ABI of libepiscan.so, the MI355X (gfx950) pair-scan engine.
 *
 * Replaces the device side of epiGPU's perform_scan: device selection and
 * queue creation (src/EpiOpenCL.c:1382-1504), the buffer uploads
 * (:1505-1545), the chunked launches of the `square` kernel (:1559-1607)
 * and the per-local-id hit slots (:1551-1555, :1606-1627).
 *
 * Plain C types only. Every call returns 0 (EPI_OK) or a negative EPI_ERR_*
 * code; epi_strerror() names the code and epi_last_error() gives the detail of
 * the last failure on a context. The library never calls exit(). A context
 * owns one HIP stream on one device; calls on a context are synchronous with
 * respect to the host unless stated and a context is not thread-safe.
 */
#ifndef EPISCAN_H
#define EPISCAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPI_OK 0
#define EPI_ERR_ARG (-1)      /* invalid argument or size */
#define EPI_ERR_HIP (-2)      /* HIP runtime error */
#define EPI_ERR_NODEV (-3)    /* no such device */
#define EPI_ERR_STATE (-4)    /* data missing for this call */
#define EPI_ERR_NOMEM (-5)    /* device allocation failed */
#define EPI_ERR_RANGE (-6)    /* phenotype cannot be encoded, bad range */

/* Hit record: same 24-byte layout as the reference's `table`
 * (src/id4117.h:14-21). snp1 > snp2, global 0-based SNP indices. */
typedef struct epi_hit {
    int32_t snp1;
    int32_t snp2;
    int32_t df1;
    int32_t df2;
    float F;
    float F2;
} epi_hit;

typedef struct epi_ctx epi_ctx;

/* Kernel forms. */
#define EPI_FORM_PLAIN 0  /* bitplanes in LDS, AND+popcount counts, FP32 sums */
#define EPI_FORM_MFMA 1   /* INT8 MFMA contraction, exact integer sums */

/* Open a context on HIP device `device` (the reference's -d,
 * src/EpiOpenCL.c:1402-1436). */
int epi_open(int device, epi_ctx **ctx);
void epi_close(epi_ctx *ctx);
const char *epi_strerror(int code);
const char *epi_last_error(const epi_ctx *ctx);
int epi_device_name(const epi_ctx *ctx, char *buf, size_t n);

/* Genotypes in layout A1 (src/readplink.c:1066-1102): nsnp x npack1 int32
 * words, npack1 = ceil(nid/16). Uploaded once and transcoded on the device
 * into class bitplanes (plain form) and class byte masks (MFMA form).
 * Replaces the per-chromosome clCreateBuffer/clEnqueueWriteBuffer of
 * src/EpiOpenCL.c:1530-1545. Resets the chromosome layout to one block over
 * all SNPs; phenotypes of a different nid are dropped. */
int epi_set_genotypes(epi_ctx *ctx, const int32_t *words, int32_t nid, int32_t nsnp);

/* nphen phenotype vectors of nid FP32 values, row-major. A value of -9 (the
 * reference's sentinel, src/EpiOpenCL.c:135) or NaN is missing and removes
 * the individual from every pair for that phenotype. Replaces the phen
 * upload and mY/SSY of src/EpiOpenCL.c:1364-1376, :1507-1508. nid must equal
 * the genotypes' nid when genotypes are set. */
int epi_set_phenotypes(epi_ctx *ctx, const float *phen, int32_t nphen, int32_t nid);

/* Phenotype flags. EPI_PHEN_NINE_IS_VALUE: -9 is an ordinary value and only
 * NaN is missing, as in the reference's safe kernels, which test only the
 * genotypes for missingness (src/EpiOpenCL.c:597-598). */
#define EPI_PHEN_NINE_IS_VALUE 1

/* As epi_set_phenotypes, with flags. */
int epi_set_phenotypes_ex(epi_ctx *ctx, const float *phen, int32_t nphen, int32_t nid,
                          int32_t flags);

/* How the MFMA form encodes one phenotype vector (host only, no device
 * needed). Replaces nothing in the reference, whose kernels take FP32 values
 * (src/EpiOpenCL.c:135-147): the MFMA form holds phenotypes as integers
 * q = round((y - centre) * 2^fbits) in three signed 8-bit limbs. The central
 * SD is that of the present values without the 64 farthest from their median
 * (a quarter of them below 258 present values). Among the 32 farthest,
 * values more than 32 central SDs out are outliers, carried exactly beside
 * the integers; the rest must lie within 64 central SDs of their centre,
 * or MFMA scans of this phenotype return EPI_ERR_RANGE (mfma_ok = 0) and the
 * plain form takes it. Returns EPI_ERR_RANGE for an infinite value or fewer
 * than two present values, as epi_set_phenotypes does. */
typedef struct epi_phen_info {
    int32_t present;            /* values present (not -9 or NaN) */
    int32_t outliers;           /* values carried exactly, 0..32 */
    int32_t mfma_ok;            /* 1: the MFMA form holds F, F2 to 1e-5 */
    int32_t fbits;              /* q = round((y - centre) * 2^fbits) */
    int32_t zbits;              /* the sum of squares uses round(q^2 / 2^zbits) */
    int32_t pad;
    double centre;              /* mean of the encoded (non-outlier) values */
    double central_sd;
    double max_dev;             /* largest |y - centre| among the encoded values */
    int32_t outlier_index[32];  /* individuals of the outliers */
} epi_phen_info;

int epi_phen_encoding(const float *phen, int32_t nid, int32_t flags, epi_phen_info *info);

/* Chromosome blocks (chromosome.chrstart / chrsize, src/readplink.h:34-39).
 * Blocks are numbered in the reference's order: block(i, j) = i(i+1)/2 + j
 * for j <= i (src/EpiOpenCL.c:1527-1540, :1445-1456). */
int epi_set_chromosomes(epi_ctx *ctx, const int32_t *chrstart, const int32_t *chrsize,
                        int32_t nchr);

typedef struct epi_scan_params {
    int32_t test;       /* 'f': 8df F > t1; 'i': (F > t1 || F2 > t2) && nfac > 6 */
    int32_t form;       /* EPI_FORM_PLAIN or EPI_FORM_MFMA */
    float t1, t2;       /* thresholds (kernel args 6-7, src/EpiOpenCL.c:1516-1517) */
    int32_t phen;       /* phenotype index */
    int32_t blk_begin;  /* first block to scan */
    int32_t blk_end;    /* one past the last block; <= 0 means all blocks */
    int32_t safe;       /* 1: the reference's safe mode (-s 1, src/EpiOpenCL.c:505-963):
                           test 'i' without the nfac > 6 condition (:949-950); load the
                           phenotypes with EPI_PHEN_NINE_IS_VALUE for its -9 handling */
    int32_t reserved[3];/* zero */
} epi_scan_params;

/* Fill defaults: test 'f', MFMA form, t1 6.5, t2 10.5 (src/main.c:244-249),
 * phenotype 0, all blocks. */
void epi_scan_params_default(epi_scan_params *p);

/* Scan every pair of the block range and collect the hits on the device.
 * Clears the previous scan's hits. */
int epi_scan(epi_ctx *ctx, const epi_scan_params *p);

/* Number of hits of the last scan (all of them, also beyond capacity) and
 * whether the device buffer overflowed. */
int epi_hit_count(epi_ctx *ctx, int64_t *nhits, int32_t *overflow);

/* Copy up to cap hits of the last scan, sorted by (block, snp1, snp2), as
 * reference records. F64/F2_64 (optional, may be NULL) receive the FP64
 * values of the same hits; block (optional) the block index of each hit.
 * *n receives the number copied. */
int epi_get_hits(epi_ctx *ctx, epi_hit *hits, double *F64, double *F2_64, int32_t *block,
                 int64_t cap, int64_t *n);

/* FP64 hit record of a device hit buffer (40 bytes): the fields of the
 * reference's `table` (src/id4117.h:14-21) with F and F2 in FP64, then the
 * phenotype index and padding. */
typedef struct epi_hit64 {
    int32_t snp1;
    int32_t snp2;
    int32_t df1;
    int32_t df2;
    double F;
    double F2;
    int32_t tag;
    int32_t pad;
} epi_hit64;

/* epi_scan without any device allocation. It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory, such as
 * tight serving loops and code that must not call hipMalloc. The tile list
 * and the hit counter live in a fixed 16 MiB device array of the library
 * (each device holds its own until the process ends), so a scan may need at
 * most 1,048,560 tiles: about 64,000 SNPs in one block without missing calls,
 * or 45,000 with them, at 32-row tiles. A larger scan fails with
 * EPI_ERR_RANGE before any device work. Genotypes and phenotypes must be set
 * (their uploads allocate). The hits, unsorted, go to the caller's device
 * buffer `hits` of `cap` records; *nhits receives their total, which may
 * exceed cap (records beyond it are not written). Synchronous; the hit
 * buffer of epi_get_hits is not touched. Calls that overlap on one device
 * share the array and race. */
int epi_scan_noalloc(epi_ctx *ctx, const epi_scan_params *p, epi_hit64 *hits, int64_t cap,
                     int64_t *nhits);

/* Device capacity of the hit buffer (records); default 1<<22. */
int epi_set_hit_capacity(epi_ctx *ctx, int64_t cap);

/* Bound on the device tile list of a single scan (tiles of 16 bytes; default
 * 1<<25, 512 MiB). A scan whose block range needs more tiles (from about
 * 260,000 to 520,000 SNPs, by tile shape) is built, uploaded and launched in
 * pieces of about this many tiles (whole rows of super-tiles), with the
 * same hits; its epi_last_scan_stats time then includes the later pieces'
 * uploads. No counterpart in the reference, which launches per chunk of
 * -i SNPs (src/EpiOpenCL.c:1564-1607). 1 <= max_tiles <= 2^30 - 1. */
int epi_set_max_tiles(epi_ctx *ctx, int64_t max_tiles);

/* Tile shape and work order of a scan. The reference's only knob here is
 * the chunk size of its launches (-i, src/EpiOpenCL.c:1559-1607); these
 * settings replace it. A FAST tile (no missing call in either SNP range) is
 * fast_rows a-SNPs x 64 b-SNPs. In the MFMA form a tile with missing calls is
 * either a GENERAL tile, fast_rows x 32 (the full {0, 1, NM} contraction), or
 * a SPARSE tile of the FAST shape (the FAST contraction plus per-pair sums
 * over the individuals with a missing call, which pay off at low missing
 * rates). A tile can be SPARSE when every SNP of its ranges misses at most
 * max(8, nid / 16) calls; dense_missing 0 makes it SPARSE when its SNPs
 * also miss at most 1/150 of their calls on average. The plain form's GENERAL
 * tiles are 16 x 64. */
typedef struct epi_schedule {
    int32_t fast_rows;  /* MFMA form: 64 or 32; plain form: 32 or 16 */
    int32_t super_a;    /* tiles are issued in super-tiles of super_a x super_b */
    int32_t super_b;    /*   tiles (1..64 each) for L2 reuse */
    int32_t queues;     /* batch only: 8 = one work queue per XCD, 1 = one shared queue */
    int32_t order;      /* batch only: 0 locality, 1 GENERAL and SPARSE tiles first,
                           2 datasets by decreasing nid */
    int32_t dense_missing; /* MFMA form, tiles with missing calls: 0 = SPARSE at
                              low missing rates (default), 1 = always GENERAL,
                              2 = SPARSE wherever allowed */
    int32_t reserved[2];
} epi_schedule;

/* Schedule of single scans of one form (default: fast_rows 32, super-tiles
 * 8 x 4). */
int epi_set_schedule(epi_ctx *ctx, int32_t form, const epi_schedule *s);
int epi_get_schedule(epi_ctx *ctx, int32_t form, epi_schedule *s);

/* Autotuner dimensions. */
#define EPI_TUNE_ROWS 1    /* fast_rows */
#define EPI_TUNE_SUPER 2   /* super_a x super_b in {8x4, 4x4, 16x2, 4x8, 1x1} */
#define EPI_TUNE_QUEUES 4  /* queues in {8, 1} */
#define EPI_TUNE_ORDER 8   /* order in {0, 1, 2} */
#define EPI_TUNE_MISSING 16 /* dense_missing in {0, 1, 2} */
#define EPI_TUNE_ALL 31

typedef struct epi_tune_entry {
    epi_schedule sched;
    double ms;          /* median device time of the scan */
} epi_tune_entry;

/* Pick the single-scan schedule for p->form by measurement: starting from
 * the current schedule, each dimension in `dims` (EPI_TUNE_ROWS and
 * EPI_TUNE_SUPER apply; 0 = both) is varied in turn and the fastest value
 * kept (coordinate descent). Every candidate runs the scan of p once as a
 * warm-up, then `reps` timed times (median). The winner becomes the
 * context's schedule for p->form. `log` (optional) receives up to log_cap
 * measured candidates, *nlog the number measured. Hit buffers hold the last
 * candidate's hits. */
int epi_autotune(epi_ctx *ctx, const epi_scan_params *p, int32_t dims, int32_t reps,
                 epi_schedule *best, double *best_ms, epi_tune_entry *log, int32_t log_cap,
                 int32_t *nlog);

/* Device time (ms, HIP events on the context's stream) of the pair kernels
 * of the last epi_scan, and the number of pairs it tested. */
int epi_last_scan_stats(epi_ctx *ctx, double *kernel_ms, int64_t *pairs, int32_t *launches);

/* Tiles of the last epi_scan, epi_scan_noalloc or epi_pair_dump by kind:
 * FAST, SPARSE and GENERAL (see epi_schedule). No counterpart in the
 * reference; for tests and measurement. */
int epi_last_tiles(epi_ctx *ctx, int64_t *fast, int64_t *sparse, int64_t *general);

/* Debug / parity entry: every pair (a, b) with a0 <= a < a1, b0 <= b < b1 and
 * a > b, through the same kernels as epi_scan, writes its 3x3 table and FP64
 * statistics into dense arrays indexed p = (a-a0)*(b1-b0) + (b-b0):
 * counts[9p+c] (c = 3*g_a + g_b), sums[9p+c] (phenotype units), F[p], F2[p]
 * (F2 computed for both tests). Entries of pairs with a <= b are untouched.
 * Any output pointer may be NULL. */
int epi_pair_dump(epi_ctx *ctx, const epi_scan_params *p, int32_t a0, int32_t a1,
                  int32_t b0, int32_t b1, int32_t *counts, double *sums, double *F,
                  double *F2);

/* ---------------------------------------------------------------- batch ---
 * Many scans in one persistent launch: datasets of different nid and nsnp,
 * each with one or more phenotypes (a scan is one (dataset, phenotype)
 * pair). Replaces a loop of perform_scan calls (src/EpiOpenCL.c:1312-1686),
 * one per dataset, per phenotype file (-P, src/main.c:172-181, :295-298) and
 * per permuted phenotype (-p, src/main.c:101-112, :309-312), with a single
 * device launch: the work items of every scan are dealt to work queues (by
 * default eight, one per XCD, of equal cost) and pulled by a grid of
 * occupancy x CU workgroups. The first epi_batch_scan with given test,
 * form and thresholds captures the launch (queue reset + scan kernel) as a
 * HIP graph; later calls with the same settings replay it.
 * A batch shares its context's stream and device; destroy it before the
 * context. */
typedef struct epi_batch epi_batch;

int epi_batch_create(epi_ctx *ctx, epi_batch **batch);
void epi_batch_destroy(epi_batch *batch);

/* Add a dataset: genotypes in layout A1 (as epi_set_genotypes) and its
 * chromosome blocks (as epi_set_chromosomes; chrstart == NULL means one block
 * over all SNPs). *dataset receives its index. */
int epi_batch_add_dataset(epi_batch *batch, const int32_t *words, int32_t nid, int32_t nsnp,
                          const int32_t *chrstart, const int32_t *chrsize, int32_t nchr,
                          int32_t *dataset);

/* Add nphen phenotype vectors of the dataset's nid (as epi_set_phenotypes);
 * each becomes one scan, numbered from *first_scan. Once per dataset. */
int epi_batch_add_phenotypes(epi_batch *batch, int32_t dataset, const float *phen,
                             int32_t nphen, int32_t *first_scan);

/* As epi_batch_add_phenotypes, with EPI_PHEN_* flags. */
int epi_batch_add_phenotypes_ex(epi_batch *batch, int32_t dataset, const float *phen,
                                int32_t nphen, int32_t flags, int32_t *first_scan);

/* Device capacity of the batch's shared hit buffer (records); default 1<<22. */
int epi_batch_set_hit_capacity(epi_batch *batch, int64_t cap);

/* Scan every pair of every block of every scan with the test, form and
 * thresholds of p (p->phen and the block range are ignored). */
int epi_batch_scan(epi_batch *batch, const epi_scan_params *p);

/* Hits of one scan (scan = -1: all scans together) of the last
 * epi_batch_scan, and whether the shared buffer overflowed. */
int epi_batch_hit_count(epi_batch *batch, int32_t scan, int64_t *nhits, int32_t *overflow);

/* As epi_get_hits, for one scan of the last epi_batch_scan. */
int epi_batch_get_hits(epi_batch *batch, int32_t scan, epi_hit *hits, double *F64,
                       double *F2_64, int32_t *block, int64_t cap, int64_t *n);

/* Device time of the last epi_batch_scan (ms, HIP events around the graph
 * launch), pairs tested over all scans, work items, graph captures made so
 * far and the persistent grid (workgroups) of the last capture. */
int epi_batch_stats(epi_batch *batch, double *kernel_ms, int64_t *pairs, int64_t *tiles,
                    int32_t *captures, int32_t *grid);

/* Schedule of the batch's scans of one form (defaults as epi_set_schedule,
 * 8 queues, locality order). */
int epi_batch_set_schedule(epi_batch *batch, int32_t form, const epi_schedule *s);
int epi_batch_get_schedule(epi_batch *batch, int32_t form, epi_schedule *s);

/* Pick the batch's schedule for p->form by measurement: starting from the
 * current schedule, each dimension in `dims` (0 = all) is varied in turn
 * and the fastest value kept (coordinate descent). Every candidate runs
 * the whole batch as a captured graph: one warm-up replay, then the median
 * of `reps` timed replays. The winner becomes the batch's schedule for
 * p->form. `log` (optional) receives up to log_cap measured candidates,
 * *nlog the number measured. Hit buffers hold the last candidate's hits. */
int epi_batch_autotune(epi_batch *batch, const epi_scan_params *p, int32_t dims, int32_t reps,
                       epi_schedule *best, double *best_ms, epi_tune_entry *log,
                       int32_t log_cap, int32_t *nlog);

#ifdef __cplusplus
}
#endif
#endif /* EPISCAN_H */
