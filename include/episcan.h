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
    int32_t reserved[4];/* zero */
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

/* Device capacity of the hit buffer (records); default 1<<22. */
int epi_set_hit_capacity(epi_ctx *ctx, int64_t cap);

/* Device time (ms, HIP events on the context's stream) of the pair kernels
 * of the last epi_scan, and the number of pairs it tested. */
int epi_last_scan_stats(epi_ctx *ctx, double *kernel_ms, int64_t *pairs, int32_t *launches);

/* Debug / parity entry: every pair (a, b) with a0 <= a < a1, b0 <= b < b1 and
 * a > b, through the same kernels as epi_scan, writes its 3x3 table and FP64
 * statistics into dense arrays indexed p = (a-a0)*(b1-b0) + (b-b0):
 * counts[9p+c] (c = 3*g_a + g_b), sums[9p+c] (phenotype units), F[p], F2[p]
 * (F2 computed for both tests). Entries of pairs with a <= b are untouched.
 * Any output pointer may be NULL. */
int epi_pair_dump(epi_ctx *ctx, const epi_scan_params *p, int32_t a0, int32_t a1,
                  int32_t b0, int32_t b1, int32_t *counts, double *sums, double *F,
                  double *F2);

#ifdef __cplusplus
}
#endif
#endif /* EPISCAN_H */
