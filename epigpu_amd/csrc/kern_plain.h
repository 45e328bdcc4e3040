// Plain popcount form of the pair scan (gfx950 VALU).
//
// Genotypes are class bitplanes (bit k of word w = individual 32w + k) laid
// out [plane][word][snp], so one stage of a SNP tile is a set of contiguous
// runs and loads coalesce. A 256-thread workgroup stages KW words (256
// individuals) of its a- and b-tiles and of the phenotype in LDS; every thread
// owns PA x PB pairs:
//   counts  X[g][h] += popcount(A_g & V & B_h)      (AND + v_bcnt per word)
//   sums    X[g][h] += y_k * [k in A_g & V] * [k in B_h]   (masked FP32 FMA)
// FP32 partial sums are promoted to FP64 after every word. The two bitplanes
// of a SNP hold its two rarest genotype classes and the most common class is
// derived (by subtraction from the NM totals), so a small cell never inherits
// the rounding error of a large one. FAST tiles use the planes {0,1}; GENERAL
// tiles add the NM plane and the sum of y^2 over complete cases. The per-pair
// epilogue is shared with the MFMA form.
#pragma once

#include "epilogue.h"
#include "kern_mfma.h"  // Marg, xcd_tile

namespace epi {

constexpr int KW = 8;  // words (of 32 individuals) per LDS stage

struct PlainScan {          // one dataset x one phenotype
    const uint32_t *bits;   // [3][Kw][nsnp]
    const float *yc;        // [Kp] centred FP32 phenotype, 0 where missing
    const uint32_t *vbits;  // [Kw] phenotype-present bits
    const Marg *marg;       // [nsnp]
    int nsnp;
    int Kw;
    int nV;
    int tag;                // written into the hits
    double ysV;             // sum of yc over present
    double y2V;             // sum of yc^2 over present
};

struct PlainArgs {
    PlainScan S;
    const int4 *tiles;
    int ntiles;
    HitSink hs;
    DumpSink ds;
};

template <int PA, int PB, bool FAST, bool DUMP>
__device__ __forceinline__ void plain_tile(const PlainScan &P, const HitSink &hs, const DumpSink &dsk, int a0, int a1,
                                           int b0, int b1)
{
    constexpr int NP = FAST ? 2 : 3;
    constexpr int TA = 16 * PA, TB = 16 * PB;
    __shared__ uint32_t sA[NP][KW][TA];
    __shared__ uint32_t sB[NP][KW][TB];
    __shared__ __attribute__((aligned(16))) float sy[KW * 32];
    __shared__ uint32_t sv[KW];

    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int nsnp = P.nsnp;

    int cnt[PA][PB][NP][NP];
    float sf[PA][PB][NP][NP];
    double sd[PA][PB][NP][NP];
    float qf[PA][PB];
    double qd[PA][PB];
#pragma unroll
    for (int i = 0; i < PA; i++)
#pragma unroll
        for (int j = 0; j < PB; j++) {
            qf[i][j] = 0.0f;
            qd[i][j] = 0.0;
#pragma unroll
            for (int g = 0; g < NP; g++)
#pragma unroll
                for (int h = 0; h < NP; h++) {
                    cnt[i][j][g][h] = 0;
                    sf[i][j][g][h] = 0.0f;
                    sd[i][j][g][h] = 0.0;
                }
        }

    constexpr int NA = NP * KW * TA, NB = NP * KW * TB, NTOT = NA + NB + KW * 32 + KW;
    for (int w0 = 0; w0 < P.Kw; w0 += KW) {
        __syncthreads();
        for (int c = tid; c < NTOT; c += 256) {
            if (c < NA) {
                int s = c % TA, w = (c / TA) % KW, p = c / (TA * KW);
                int snp = a0 + s < nsnp ? a0 + s : nsnp - 1;
                sA[p][w][s] = P.bits[((size_t)p * P.Kw + w0 + w) * nsnp + snp];
            } else if (c < NA + NB) {
                int d = c - NA;
                int s = d % TB, w = (d / TB) % KW, p = d / (TB * KW);
                int snp = b0 + s < nsnp ? b0 + s : nsnp - 1;
                sB[p][w][s] = P.bits[((size_t)p * P.Kw + w0 + w) * nsnp + snp];
            } else if (c < NA + NB + KW * 32) {
                int d = c - NA - NB;
                sy[d] = P.yc[(size_t)w0 * 32 + d];
            } else {
                int d = c - NA - NB - KW * 32;
                sv[d] = P.vbits[w0 + d];
            }
        }
        __syncthreads();

        for (int w = 0; w < KW; w++) {
            const uint32_t v = sv[w];
            uint32_t aw[PA][NP], bw[PB][NP];
#pragma unroll
            for (int i = 0; i < PA; i++)
#pragma unroll
                for (int g = 0; g < NP; g++) aw[i][g] = sA[g][w][ty + 16 * i] & v;
#pragma unroll
            for (int j = 0; j < PB; j++)
#pragma unroll
                for (int h = 0; h < NP; h++) bw[j][h] = sB[h][w][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < PA; i++)
#pragma unroll
                for (int j = 0; j < PB; j++)
#pragma unroll
                    for (int g = 0; g < NP; g++)
#pragma unroll
                        for (int h = 0; h < NP; h++) cnt[i][j][g][h] += __popc(aw[i][g] & bw[j][h]);
            const float4 *y4 = reinterpret_cast<const float4 *>(sy + 32 * w);
#pragma unroll
            for (int k4 = 0; k4 < 8; k4++) {
                const float4 yv = y4[k4];
                const float yk[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const int k = 4 * k4 + kk;
                    const int ybits = __float_as_int(yk[kk]);
                    float ya[PA][NP], bb[PB][NP];
#pragma unroll
                    for (int i = 0; i < PA; i++)
#pragma unroll
                        for (int g = 0; g < NP; g++)
                            ya[i][g] = __int_as_float(__builtin_amdgcn_sbfe((int)aw[i][g], k, 1) & ybits);
#pragma unroll
                    for (int j = 0; j < PB; j++)
#pragma unroll
                        for (int h = 0; h < NP; h++)
                            bb[j][h] = (float)__builtin_amdgcn_ubfe(bw[j][h], k, 1);
#pragma unroll
                    for (int i = 0; i < PA; i++)
#pragma unroll
                        for (int j = 0; j < PB; j++)
#pragma unroll
                            for (int g = 0; g < NP; g++)
#pragma unroll
                                for (int h = 0; h < NP; h++) sf[i][j][g][h] = fmaf(ya[i][g], bb[j][h], sf[i][j][g][h]);
                    if (!FAST) {
#pragma unroll
                        for (int i = 0; i < PA; i++) {
                            const float y2 = ya[i][2] * yk[kk];
#pragma unroll
                            for (int j = 0; j < PB; j++) qf[i][j] = fmaf(y2, bb[j][2], qf[i][j]);
                        }
                    }
                }
            }
            // promote the word's FP32 partial sums
#pragma unroll
            for (int i = 0; i < PA; i++)
#pragma unroll
                for (int j = 0; j < PB; j++) {
                    qd[i][j] += (double)qf[i][j];
                    qf[i][j] = 0.0f;
#pragma unroll
                    for (int g = 0; g < NP; g++)
#pragma unroll
                        for (int h = 0; h < NP; h++) {
                            sd[i][j][g][h] += (double)sf[i][j][g][h];
                            sf[i][j][g][h] = 0.0f;
                        }
                }
        }
    }

#pragma unroll
    for (int i = 0; i < PA; i++) {
        const int a = a0 + ty + 16 * i;
        const Marg ma = P.marg[a < nsnp ? a : nsnp - 1];
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int b = b0 + tx + 16 * j;
            const Marg mb = P.marg[b < nsnp ? b : nsnp - 1];
            int xc[3][3];
            double xs[3][3];
            double sq;
            if (FAST) {
#pragma unroll
                for (int g = 0; g < 2; g++) {
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        xc[g][h] = cnt[i][j][g][h];
                        xs[g][h] = sd[i][j][g][h];
                    }
                    xc[g][2] = ma.pn[g];
                    xs[g][2] = ma.ys[g];
                    xc[2][g] = mb.pn[g];
                    xs[2][g] = mb.ys[g];
                }
                xc[2][2] = P.nV;
                xs[2][2] = P.ysV;
                sq = P.y2V;
            } else {
#pragma unroll
                for (int g = 0; g < 3; g++)
#pragma unroll
                    for (int h = 0; h < 3; h++) {
                        xc[g][h] = cnt[i][j][g % NP][h % NP];
                        xs[g][h] = sd[i][j][g % NP][h % NP];
                    }
                sq = qd[i][j];
            }
            int n[9];
            double s[9];
            float sf[9];
            cells_from_x(xc, n);
            cells_from_x(xs, s);
#pragma unroll
            for (int c = 0; c < 9; c++) sf[c] = (float)s[c];
            auto exact = [&](double sd[9], double &sqd) {
#pragma unroll
                for (int c = 0; c < 9; c++) sd[c] = s[c];
                sqd = sq;
            };
            const bool valid = a < a1 && b < b1 && a > b;
            finish_pair<DUMP>(hs, dsk, valid, a, b, n, sf, (float)sq, exact, ma.cls, mb.cls);
        }
    }
}

template <int PA, int PB, bool FAST, bool DUMP>
__global__ __launch_bounds__(256) void plain_kernel(const PlainArgs A)
{
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    HitSink hs = A.hs;
    hs.tag = A.S.tag;
    plain_tile<PA, PB, FAST, DUMP>(A.S, hs, A.ds, T.x, T.y, T.z, T.w);
}

}  // namespace epi
