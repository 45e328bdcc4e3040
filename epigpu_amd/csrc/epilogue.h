// Per-pair epilogue shared by the plain and MFMA pair kernels.
//
// Both kernels deliver, per SNP pair (a, b), the table X over the classes
// {0, 1, NM} x {0, 1, NM} (NM = genotype present) of complete-case counts and
// phenotype sums, plus the complete-case sum of squares. The nine cells
// c = 3*g_a + g_b of the reference (src/EpiOpenCL.c:140) follow from X by
// inclusion-exclusion, e.g. n[g][2] = X[g][NM] - X[g][0] - X[g][1].
//
// Statistics follow kernelint / kernelfull (src/EpiOpenCL.c:455-486, :231-247):
//   SSB = sum n_c (m_c - mbar)^2, SSW = SQ - N' mbar^2 - SSB,
//   F   = (SSB / (nfac-1)) / (SSW / (N'-nfac)),
//   SSI = sum n_c (m_c - R_ga - C_gb + mbar)^2, F2 = (SSI/4) / (SSW / d2).
// A pair is first screened in FP32 against thresholds lowered by 2%; pairs
// that pass are recomputed in FP64 and decided exactly, so the hit set and the
// reported F/F2 are FP64 results. The hit rules are the reference's
// (src/EpiOpenCL.c:251, :487-489); F is undefined and never a hit unless
// d1 >= 1 and d2 >= 1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace epi {

struct DevHit {
    int32_t snp1, snp2, df1, df2;
    double F, F2;
    int32_t tag, pad;  // tag: phenotype (single scan) or scan id (batch)
};
static_assert(sizeof(DevHit) == 40, "DevHit is the 40-byte epi_hit64 record");

struct HitSink {
    DevHit *hits;
    unsigned long long *nhits;
    long long cap;
    int32_t test;
    float t1, t2;    // exact thresholds
    float t1s, t2s;  // screening thresholds (lowered)
    int32_t tag;     // written into every hit
    int32_t safe;    // 1: test 'i' without the nfac > 6 condition (the reference's safe kernels)
};

struct DumpSink {
    int32_t *counts;
    double *sums;
    double *F;
    double *F2;
    int32_t a0, a1, b0, b1;
    double unit;    // phenotype units per kernel sum unit
    double centre;  // phenotype shift removed before the kernel
};

// Cells from the {0,1,NM}^2 table (counts).
template <class T>
__device__ __forceinline__ void cells_from_x(const T x[3][3], T c[9])
{
    c[0] = x[0][0];
    c[1] = x[0][1];
    c[3] = x[1][0];
    c[4] = x[1][1];
    c[2] = x[0][2] - x[0][0] - x[0][1];
    c[5] = x[1][2] - x[1][0] - x[1][1];
    c[6] = x[2][0] - x[0][0] - x[1][0];
    c[7] = x[2][1] - x[0][1] - x[1][1];
    c[8] = x[2][2] - x[0][2] - x[1][2] - x[2][0] - x[2][1] + x[0][0] + x[0][1] + x[1][0] + x[1][1];
}

// FP64 statistics. s[] and sq in any consistent phenotype units.
template <class S>
__device__ __forceinline__ void stats64(const int n[9], const S s[9], double sq, double &F, double &F2,
                                        int &d1, int &d2, int &nfac)
{
    int N = 0;
    double tot = 0.0;
    nfac = 0;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        N += n[c];
        tot += (double)s[c];
        nfac += n[c] > 0;
    }
    d1 = nfac - 1;
    d2 = N - nfac;
    if (N <= 0) {
        F = F2 = __builtin_nan("");
        return;
    }
    double mbar = tot / (double)N, m[9];
#pragma unroll
    for (int c = 0; c < 9; c++) m[c] = n[c] > 0 ? (double)s[c] / (double)n[c] : 0.0;
    double R[3], C[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        int nr = n[3 * r] + n[3 * r + 1] + n[3 * r + 2];
        int nc = n[r] + n[r + 3] + n[r + 6];
        double sr = (double)s[3 * r] + (double)s[3 * r + 1] + (double)s[3 * r + 2];
        double sc = (double)s[r] + (double)s[r + 3] + (double)s[r + 6];
        R[r] = nr > 0 ? sr / (double)nr : 0.0;
        C[r] = nc > 0 ? sc / (double)nc : 0.0;
    }
    double ssb = 0.0, ssi = 0.0;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        if (n[c] > 0) {
            double e = m[c] - mbar;
            ssb += (double)n[c] * e * e;
            double t = m[c] - R[c / 3] - C[c % 3] + mbar;
            ssi += (double)n[c] * t * t;
        }
    }
    double ssw = sq - (double)N * mbar * mbar - ssb;
    double msw = ssw / (double)d2;
    F = (ssb / (double)d1) / msw;
    F2 = (ssi * 0.25) / msw;
}

__device__ __forceinline__ bool hit_rule(int test, int safe, double F, double F2, int nfac, int d1, int d2,
                                         float t1, float t2)
{
    if (d1 < 1 || d2 < 1) return false;
    if (test == 'f') return F > (double)t1;
    return (F > (double)t1 || F2 > (double)t2) && (safe || nfac > 6);
}

// FP32 screen: false only if the pair certainly misses both thresholds.
__device__ __forceinline__ bool screen32(const int n[9], const float s[9], float sq, int test, int safe,
                                         float t1s, float t2s)
{
    int N = 0, nfac = 0;
    float tot = 0.0f;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        N += n[c];
        tot += s[c];
        nfac += n[c] > 0;
    }
    int d1 = nfac - 1, d2 = N - nfac;
    if (d1 < 1 || d2 < 1) return false;
    if (test != 'f' && !safe && nfac <= 6) return false;
    float rN = __builtin_amdgcn_rcpf((float)N);
    float mbar = tot * rN;
    float m[9];
#pragma unroll
    for (int c = 0; c < 9; c++) m[c] = n[c] > 0 ? s[c] * __builtin_amdgcn_rcpf((float)n[c]) : 0.0f;
    float ssb = 0.0f;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        float e = m[c] - mbar;
        ssb += (float)n[c] * e * e;
    }
    float ssw = sq - (float)N * mbar * mbar - ssb;
    // degenerate, or SSW a small difference of large terms (an outlier alone in its cell): FP32 cannot
    // decide within the 2% margin, so FP64 does
    if (!(ssw > 1e-4f * (sq + (float)N * mbar * mbar + ssb))) return true;
    // F > t  <=>  ssb * d2 > t * d1 * ssw
    if (ssb * (float)d2 > t1s * (float)d1 * ssw) return true;
    if (test == 'f') return false;
    // SSI <= SSB: with e the cell deviations m_c - mbar and <x,y> = sum n_c x_c y_c, the
    // row and column terms are the projections Pr e, Pc e onto row- and column-constant
    // vectors, and SSI = |e - Pr e - Pc e|^2 = |e|^2 - |Pr e|^2 - |Pc e|^2 + 2 <Pr e, Pc e>
    // <= |e|^2 = SSB. Hence F2 <= d1 F / 4, and no F2 hit is possible unless
    // SSB d2 / 4 > t2 SSW, which almost no pair meets: skip the F2 terms.
    if (!(0.25f * ssb * (float)d2 > t2s * ssw)) return false;
    float R[3], C[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        int nr = n[3 * r] + n[3 * r + 1] + n[3 * r + 2];
        int nc = n[r] + n[r + 3] + n[r + 6];
        R[r] = nr > 0 ? (s[3 * r] + s[3 * r + 1] + s[3 * r + 2]) * __builtin_amdgcn_rcpf((float)nr) : 0.0f;
        C[r] = nc > 0 ? (s[r] + s[r + 3] + s[r + 6]) * __builtin_amdgcn_rcpf((float)nc) : 0.0f;
    }
    float ssi = 0.0f;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        if (n[c] > 0) {
            float t = m[c] - R[c / 3] - C[c % 3] + mbar;
            ssi += (float)n[c] * t * t;
        }
    }
    return 0.25f * ssi * (float)d2 > t2s * ssw;
}

// First stage of the screen for FAST tiles, where every present individual is
// a complete case, so that N, the phenotype total T and SST = sq - T^2 / N are
// the same for every pair: SSB = sum_c s_c^2 / n_c - T^2 / N needs one
// reciprocal and two multiply-adds per cell. n are the counts as FP32 (exact
// below 2^24). False only if screen32 is false for certain (the same rules and
// the same 2% margin; FP32 rounding moves SSB by far less than the margin,
// and an SSW below 1e-4 of the terms it is the difference of defers to FP64,
// as in screen32).
__device__ __forceinline__ bool screen_fast_maybe(const float n[9], const float s[9], float N, float T2N, float sq,
                                                  int test, int safe, float t1s, float t2s)
{
    float nfac = 0.0f, ssb = -T2N;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        const bool occ = n[c] > 0.0f;
        nfac += occ ? 1.0f : 0.0f;
        ssb += occ ? s[c] * s[c] * __builtin_amdgcn_rcpf(n[c]) : 0.0f;
    }
    const float d1 = nfac - 1.0f, d2 = N - nfac;
    const bool dfok = d1 >= 1.0f && d2 >= 1.0f && (test == 'f' || safe || nfac > 6.0f);
    const float ssw = sq - T2N - ssb;
    const bool guard = !(ssw > 1e-4f * (sq + T2N + ssb));
    const bool f1 = ssb * d2 > t1s * d1 * ssw;
    const bool f2 = test != 'f' && 0.25f * ssb * d2 > t2s * ssw;
    return dfok && (guard || f1 || f2);
}

// screen_fast_maybe without per-cell selects: nfac = sum min(n_c, 1), and
// an empty cell's term is s_c^2 / max(n_c, 1), where s_c is zero but for FP32
// rounding of the inclusion-exclusion; such a residue only raises SSB (and
// lowers SSW), so the result is true wherever screen_fast_maybe's is.
__device__ __forceinline__ bool screen_fast_lane(const float n[9], const float s[9], float N, float T2N, float sq,
                                                 int test, int safe, float t1s, float t2s)
{
    float nfac = 0.0f, ssb = -T2N;
#pragma unroll
    for (int c = 0; c < 9; c++) {
        nfac += fminf(n[c], 1.0f);
        ssb = __builtin_fmaf(s[c] * s[c], __builtin_amdgcn_rcpf(fmaxf(n[c], 1.0f)), ssb);
    }
    const float d1 = nfac - 1.0f, d2 = N - nfac;
    const bool dfok = (d1 >= 1.0f) & (d2 >= 1.0f) & ((test == 'f') | (safe != 0) | (nfac > 6.0f));
    const float ssw = sq - T2N - ssb;
    const bool guard = !(ssw > 1e-4f * (sq + T2N + ssb));
    const bool f1 = ssb * d2 > t1s * d1 * ssw;
    const bool f2 = (test != 'f') & (0.25f * ssb * d2 > t2s * ssw);
    return dfok & (guard | f1 | f2);
}

// screen_fast_maybe for a pair whose nine cells are all occupied (nfac = 9,
// d1 = 8, d2 = N - 9 >= 1, which the caller checks): with c1 = 8 t1s / d2 and
// c2 = 4 t2s / d2 (c2 < 0 for test f, so that its test never holds), F > t1s
// is ssb d2 > 8 t1s ssw, i.e. ssb > c1 ssw (d2 > 0), and the F2 bound is ssb >
// c2 ssw; SST = sq - T^2 / N. The same decisions as screen_fast_maybe.
__device__ __forceinline__ bool screen_fast9(const float n[9], const float s[9], float T2N, float SST, float sq,
                                             float c1, float c2)
{
    float ssb = -T2N;
#pragma unroll
    for (int c = 0; c < 9; c++) ssb = __builtin_fmaf(s[c] * s[c], __builtin_amdgcn_rcpf(n[c]), ssb);
    const float ssw = SST - ssb;
    const bool guard = !(ssw > 1e-4f * (sq + T2N + ssb));
    return guard | (ssb > c1 * ssw) | ((c2 >= 0.0f) & (ssb > c2 * ssw));
}

// Wave-aggregated append: one atomic per wave and call. Every lane of the
// wave must call it (hit may be false).
__device__ __forceinline__ void append_hit(const HitSink &hs, bool hit, const DevHit &h)
{
    unsigned long long mask = __ballot(hit);
    if (mask == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(hs.nhits, (unsigned long long)__popcll(mask));
    unsigned lo = __shfl((unsigned)(base & 0xffffffffull), leader);
    unsigned hi = __shfl((unsigned)(base >> 32), leader);
    base = ((unsigned long long)hi << 32) | lo;
    if (hit) {
        unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
        unsigned long long slot = base + (unsigned long long)__popcll(below);
        if (slot < (unsigned long long)hs.cap) hs.hits[slot] = h;
    }
}

// FP64 decision for a pair that passed the screen.
__device__ __forceinline__ bool confirm64(const HitSink &hs, int a, int b, const int *n, const double *s, double sq,
                                       DevHit &h)
{
    double F, F2;
    int d1, d2, nfac;
    stats64(n, s, sq, F, F2, d1, d2, nfac);
    h.snp1 = a;
    h.snp2 = b;
    h.df1 = d1;
    h.df2 = d2;
    h.F = F;
    h.F2 = hs.test == 'f' ? 0.0 : F2;
    h.tag = hs.tag;
    h.pad = 0;
    return hit_rule(hs.test, hs.safe, F, F2, nfac, d1, d2, hs.t1, hs.t2);
}

// Finish one pair. n[9] are the exact cell counts; sf[9] and sqf are FP32
// cell sums and sum of squares, good enough to screen; exact(sd, sq) fills
// the exact sums (FP64, same units) and is called only for pairs that pass the
// screen, or for every pair when dumping. Cells are indexed 3*pa + pb by
// plane; clsa/clsb map plane -> genotype class (2 bits per plane): the
// statistics do not depend on the labels, only the dump writes cells back in
// genotype order. valid = the pair is inside the tile and a > b. All lanes of
// the wave must call.
template <bool DUMP, class Exact>
__device__ __forceinline__ void finish_pair(const HitSink &hs, const DumpSink &ds, bool valid, int a, int b,
                                            const int n[9], const float sf[9], float sqf, Exact exact, int clsa,
                                            int clsb)
{
    if (DUMP) {
        if (valid && a >= ds.a0 && a < ds.a1 && b >= ds.b0 && b < ds.b1) {
            double sd[9], sq, F, F2;
            int d1, d2, nfac;
            exact(sd, sq);
            stats64(n, sd, sq, F, F2, d1, d2, nfac);
            long long p = (long long)(a - ds.a0) * (ds.b1 - ds.b0) + (b - ds.b0);
            for (int c = 0; c < 9; c++) {
                int o = 3 * ((clsa >> (2 * (c / 3))) & 3) + ((clsb >> (2 * (c % 3))) & 3);
                if (ds.counts) ds.counts[9 * p + o] = n[c];
                if (ds.sums) ds.sums[9 * p + o] = sd[c] * ds.unit + (double)n[c] * ds.centre;
            }
            if (ds.F) ds.F[p] = F;
            if (ds.F2) ds.F2[p] = F2;
        }
        return;
    }
    const bool cand = valid && screen32(n, sf, sqf, hs.test, hs.safe, hs.t1s, hs.t2s);
    bool hit = false;
    DevHit h;
    if (cand) {
        double sd[9], sq;
        exact(sd, sq);
        hit = confirm64(hs, a, b, n, sd, sq, h);
    }
    append_hit(hs, hit, h);
    // keep the scheduler from interleaving the temporaries of several pairs
    __builtin_amdgcn_sched_barrier(0);
}

}  // namespace epi
