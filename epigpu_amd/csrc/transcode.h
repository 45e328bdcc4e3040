// Per-dataset preparation kernels: the 2-bit layout A1 (16 genotypes per
// int32, first individual in the most significant bits, right-aligned tail
// word) is decoded once into the operand layouts of the two pair-kernel forms,
// and per-SNP marginals are computed per phenotype.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kern_mfma.h"

namespace epi {

__device__ __forceinline__ int decode(const uint32_t *w, int npack, int remain, int nid, int k)
{
    if (k >= nid) return -1;
    int j = k >> 4, t = k & 15;
    int sh = j < npack ? 30 - 2 * t : 2 * (remain - 1 - t);
    return (int)((w[j] >> sh) & 3u);
}

// One thread per SNP: order the genotype classes by frequency and count the
// missing calls. The plain form stores the two rarest classes as bitplanes 0
// and 1 and derives the most common one; cls packs plane -> class (2 bits
// each).
__global__ void class_order(const uint32_t *words, int nid, int npack, int remain, int npack1, int nsnp,
                            int32_t *cls, int32_t *miss)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsnp) return;
    const uint32_t *w = words + (size_t)s * npack1;
    int n[4] = {0, 0, 0, 0};
    for (int k = 0; k < nid; k++) n[decode(w, npack, remain, nid, k)]++;
    int big = 0;
    if (n[1] > n[big]) big = 1;
    if (n[2] > n[big]) big = 2;
    int p0 = big == 0 ? 1 : 0;
    int p1 = big == 2 ? 1 : 2;
    cls[s] = p0 | (p1 << 2) | (big << 4);
    miss[s] = n[3];
}

// One thread per (word of 32 individuals, snp): bitplanes [3][Kw][nsnp] of
// the classes cls names for planes 0 and 1 (cls == nullptr: classes 0 and 1,
// the MFMA form's planes) and of the call-present flag in plane 2.
__global__ void transcode_bits(const uint32_t *words, int nid, int npack, int remain, int npack1, int nsnp,
                               int Kw, const int32_t *cls, uint32_t *bits)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int wd = blockIdx.y;
    if (s >= nsnp) return;
    const uint32_t *w = words + (size_t)s * npack1;
    const int c0 = cls ? cls[s] & 3 : 0, c1 = cls ? (cls[s] >> 2) & 3 : 1;
    uint32_t b0 = 0, b1 = 0, bn = 0;
    for (int t = 0; t < 32; t++) {
        int c = decode(w, npack, remain, nid, 32 * wd + t);
        b0 |= (uint32_t)(c == c0) << t;
        b1 |= (uint32_t)(c == c1) << t;
        bn |= (uint32_t)(c >= 0 && c != 3) << t;
    }
    bits[((size_t)0 * Kw + wd) * nsnp + s] = b0;
    bits[((size_t)1 * Kw + wd) * nsnp + s] = b1;
    bits[((size_t)2 * Kw + wd) * nsnp + s] = bn;
}

// Per SNP, over individuals with a call and a present phenotype: counts and
// phenotype sums of genotype classes 0, 1 and of all calls (NM), stored per
// class for the MFMA form and per bitplane (cls order) for the plain form.
// One workgroup per 64 SNPs: lane = SNP (coalesced bitplane loads), four
// groups of threads split the words and are reduced in LDS.
__global__ __launch_bounds__(256) void marginals_kernel(const uint32_t *mbits, int Kw, int nsnp,
                                                        const uint32_t *vbits, const int32_t *qv,
                                                        const float *yc, const int32_t *cls, const MissPhen *mp,
                                                        Marg *marg)
{
    __shared__ long long sq[4][3][64];
    __shared__ double sy[4][3][64];
    __shared__ int sn[4][3][64];
    __shared__ long long sz[4][64];
    const int sl = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + sl;
    int n[3] = {0, 0, 0};
    long long q[3] = {0, 0, 0}, z = 0;
    double y[3] = {0.0, 0.0, 0.0};
    if (s < nsnp) {
        for (int wd = grp; wd < Kw; wd += 4) {
            const uint32_t v = vbits[wd];
            const uint32_t b0 = mbits[((size_t)0 * Kw + wd) * nsnp + s] & v;
            const uint32_t b1 = mbits[((size_t)1 * Kw + wd) * nsnp + s] & v;
            const uint32_t bn = mbits[((size_t)2 * Kw + wd) * nsnp + s] & v;
            const uint32_t m[3] = {b0, b1, bn & ~b0 & ~b1};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                uint32_t x = m[c];
                n[c] += __popc(x);
                while (x) {
                    const int k = 32 * wd + __ffs(x) - 1;
                    x &= x - 1;
                    q[c] += qv[k];
                    y[c] += (double)yc[k];
                    z += mp[k].z;
                }
            }
        }
    }
    for (int c = 0; c < 3; c++) {
        sn[grp][c][sl] = n[c];
        sq[grp][c][sl] = q[c];
        sy[grp][c][sl] = y[c];
    }
    sz[grp][sl] = z;
    __syncthreads();
    if (grp == 0 && s < nsnp) {
        int tn[3];
        long long tq[3];
        double ty[3];
        for (int c = 0; c < 3; c++) {
            tn[c] = sn[0][c][sl] + sn[1][c][sl] + sn[2][c][sl] + sn[3][c][sl];
            tq[c] = sq[0][c][sl] + sq[1][c][sl] + sq[2][c][sl] + sq[3][c][sl];
            ty[c] = ((sy[0][c][sl] + sy[1][c][sl]) + sy[2][c][sl]) + sy[3][c][sl];
        }
        Marg mg;
        const int cl = cls[s];
        const int ntot = tn[0] + tn[1] + tn[2];
        mg.n[0] = tn[0];
        mg.n[1] = tn[1];
        mg.n[2] = ntot;
        mg.q[0] = tq[0];
        mg.q[1] = tq[1];
        mg.q[2] = tq[0] + tq[1] + tq[2];
        for (int p = 0; p < 2; p++) {
            int c = (cl >> (2 * p)) & 3;
            mg.pn[p] = tn[c];
            mg.ys[p] = ty[c];
        }
        mg.pn[2] = ntot;
        mg.ys[2] = ty[0] + ty[1] + ty[2];
        mg.cls = cl;
        mg.pad = 0;
        mg.zs = sz[0][sl] + sz[1][sl] + sz[2][sl] + sz[3][sl];
        marg[s] = mg;
    }
}

// Codes of one phenotype's outlier individuals per SNP (thread = SNP): two
// bits per outlier, the genotype class (0, 1, 2) or 3 for a missing call;
// word w holds outliers 16w..16w+15, stored [w][snp] so that a wave's SNPs
// read consecutive words.
__global__ __launch_bounds__(256) void outlier_codes_kernel(const uint32_t *mbits, int Kw, int nsnp,
                                                            const int32_t *oidx, int nout, uint32_t *ocodes)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsnp) return;
    for (int w = 0; 16 * w < nout; w++) {
        uint32_t code = 0;
        for (int t = 0; t < 16 && 16 * w + t < nout; t++) {
            const int k = oidx[16 * w + t], wd = k >> 5;
            const uint32_t bit = 1u << (k & 31);
            const bool c0 = (mbits[((size_t)0 * Kw + wd) * nsnp + s] & bit) != 0;
            const bool c1 = (mbits[((size_t)1 * Kw + wd) * nsnp + s] & bit) != 0;
            const bool nm = (mbits[((size_t)2 * Kw + wd) * nsnp + s] & bit) != 0;
            const uint32_t g = !nm ? 3u : c0 ? 0u : c1 ? 1u : 2u;
            code |= g << (2 * t);
        }
        ocodes[(size_t)w * nsnp + s] = code;
    }
}

// One thread per SNP: the individuals (below nid) without a call, ascending,
// into mlist[moff[s] ...].
__global__ __launch_bounds__(256) void miss_list_kernel(const uint32_t *mbits, int Kw, int nsnp, int nid,
                                                        const long long *moff, int32_t *mlist)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsnp) return;
    long long o = moff[s];
    for (int wd = 0; wd < Kw && 32 * wd < nid; wd++) {
        uint32_t x = ~mbits[((size_t)2 * Kw + wd) * nsnp + s];
        if (nid - 32 * wd < 32) x &= (1u << (nid - 32 * wd)) - 1u;
        while (x) {
            mlist[o++] = 32 * wd + __ffs(x) - 1;
            x &= x - 1u;
        }
    }
}

// Transposed bitplanes for the SPARSE tiles: one wave per (64-SNP word wv of
// the virtual SNP order, 32-individual word wd). Lane l is virtual SNP
// 64 wv + l (snp_of_v: its SNP, or -1 for padding); a ballot over the lanes
// of bit k of a plane word is the 64-SNP word of individual 32 wd + k, which
// lane k collects and stores: tbits[wv][32 wd + k][plane], planes class 0,
// class 1, call present, and a zero word (a 64-SNP column is contiguous).
__global__ __launch_bounds__(64) void tbits_kernel(const uint32_t *mbits, int Kw, int nsnp, int nid,
                                                    const int32_t *snp_of_v, int nw, uint64_t *tbits)
{
    const int wv = blockIdx.x, wd = blockIdx.y, l = threadIdx.x;
    const int s = snp_of_v[64 * wv + l];
    uint32_t lo[3] = {0u, 0u, 0u}, hi[3] = {0u, 0u, 0u};
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const uint32_t x = s >= 0 ? mbits[((size_t)p * Kw + wd) * nsnp + s] : 0u;
        for (int k = 0; k < 32; k++) {
            const uint64_t b = __ballot((x >> k) & 1u);
            lo[p] = l == k ? (uint32_t)b : lo[p];
            hi[p] = l == k ? (uint32_t)(b >> 32) : hi[p];
        }
    }
    const int i = 32 * wd + l;
    if (l < 32 && i < nid) {
        uint64_t *o = tbits + ((size_t)wv * nid + i) * 4;
#pragma unroll
        for (int p = 0; p < 3; p++) o[p] = ((uint64_t)hi[p] << 32) | lo[p];
        o[3] = 0ull;
    }
}

}  // namespace epi
