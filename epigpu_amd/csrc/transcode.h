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

// One thread per (snp, 16-individual group): class masks for the MFMA form.
__global__ void transcode_planes(const uint32_t *words, int nid, int npack, int remain, int npack1,
                                 int nsnp, int Kp, uint8_t *planes, int32_t *miss)
{
    const int groups = Kp / 16;
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)groups * nsnp) return;
    int s = (int)(gid / groups), grp = (int)(gid % groups);
    const uint32_t *w = words + (size_t)s * npack1;
    uint32_t m[3][4] = {{0}};
    int anymiss = 0;
    for (int t = 0; t < 16; t++) {
        int c = decode(w, npack, remain, nid, 16 * grp + t);
        uint32_t byte = 0xFFu << (8 * (t & 3));
        if (c == 0) m[0][t >> 2] |= byte;
        if (c == 1) m[1][t >> 2] |= byte;
        if (c >= 0 && c != 3) m[2][t >> 2] |= byte;
        anymiss |= c == 3;
    }
    for (int p = 0; p < 3; p++)
        *reinterpret_cast<uint4 *>(planes + ((size_t)s * 3 + p) * Kp + 16 * grp) =
            make_uint4(m[p][0], m[p][1], m[p][2], m[p][3]);
    if (anymiss) atomicOr(&miss[s], 1);
}

// One thread per SNP: order the genotype classes by frequency. The plain form
// stores the two rarest classes as bitplanes 0 and 1 and derives the most
// common one; cls packs plane -> class (2 bits each).
__global__ void class_order(const uint32_t *words, int nid, int npack, int remain, int npack1, int nsnp,
                            int32_t *cls)
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
}

// One thread per (word of 32 individuals, snp): class bitplanes, plain form.
__global__ void transcode_bits(const uint32_t *words, int nid, int npack, int remain, int npack1, int nsnp,
                               int Kw, const int32_t *cls, uint32_t *bits)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int wd = blockIdx.y;
    if (s >= nsnp) return;
    const uint32_t *w = words + (size_t)s * npack1;
    const int c0 = cls[s] & 3, c1 = (cls[s] >> 2) & 3;
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

// One workgroup per SNP: per genotype class (0, 1, 2) counts and phenotype
// sums over present phenotypes, for one phenotype; stored per class for the
// MFMA form and per bitplane (cls order) for the plain form.
__global__ __launch_bounds__(256) void marginals_kernel(const uint8_t *planes, int Kp, int nsnp,
                                                        const int8_t *wts, const int32_t *qv,
                                                        const float *yc, const int32_t *cls, Marg *marg)
{
    __shared__ long long sq[3][256];
    __shared__ double sy[3][256];
    __shared__ int sn[3][256];
    const int s = blockIdx.x, tid = threadIdx.x;
    int n[3] = {0, 0, 0};
    long long q[3] = {0, 0, 0};
    double y[3] = {0.0, 0.0, 0.0};
    const uint8_t *pl = planes + (size_t)s * 3 * Kp;
    for (int k = tid; k < Kp; k += 256) {
        if (!wts[k]) continue;  // row 0 = V
        if (!pl[(size_t)2 * Kp + k]) continue;  // missing call
        int c = pl[k] ? 0 : (pl[(size_t)Kp + k] ? 1 : 2);
        n[c] += 1;
        q[c] += qv[k];
        y[c] += (double)yc[k];
    }
    for (int p = 0; p < 3; p++) {
        sn[p][tid] = n[p];
        sq[p][tid] = q[p];
        sy[p][tid] = y[p];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int p = 0; p < 3; p++) {
                sn[p][tid] += sn[p][tid + st];
                sq[p][tid] += sq[p][tid + st];
                sy[p][tid] += sy[p][tid + st];
            }
        __syncthreads();
    }
    if (tid == 0) {
        Marg m;
        const int cl = cls[s];
        const int ntot = sn[0][0] + sn[1][0] + sn[2][0];
        m.n[0] = sn[0][0];
        m.n[1] = sn[1][0];
        m.n[2] = ntot;
        m.q[0] = sq[0][0];
        m.q[1] = sq[1][0];
        m.q[2] = sq[0][0] + sq[1][0] + sq[2][0];
        for (int p = 0; p < 2; p++) {
            int c = (cl >> (2 * p)) & 3;
            m.pn[p] = sn[c][0];
            m.ys[p] = sy[c][0];
        }
        m.pn[2] = ntot;
        m.ys[2] = sy[0][0] + sy[1][0] + sy[2][0];
        m.cls = cl;
        m.pad = 0;
        marg[s] = m;
    }
}

}  // namespace epi
