// MFMA contraction form of the pair scan (gfx950, v_mfma_i32_32x32x32_i8).
//
// For a tile of a-SNPs x b-SNPs the cell table is one INT8 contraction over
// individuals k:
//     X[g][h][w](a, b) = sum_k  (I_g(a,k) * W_w(k)) * I_h(b,k)
// where I_g are 0/1 class indicators (g in {0, 1, NM}) and W_w the per-
// individual weights: W_0 = V (phenotype present, 0/1), W_1..3 = the three
// signed 8-bit limbs of the fixed-point phenotype q = L0 + 256 L1 + 65536 L2,
// W_4..6 the limbs of round(q^2 / 2^23). Every product and every INT32 sum is
// exact, so counts are bit-exact and cell sums are exact integers in units of
// the fixed-point phenotype.
//
// Operands: class masks are stored as bytes 0xFF / 0x00 (int8 -1 / 0) in HBM,
// [snp][plane][Kp], and staged per 128-individual chunk into LDS. An A operand
// row is (a, g, w): mask(a, g) AND W_w. A B operand is the raw mask of 32
// b-SNPs for one class, so every product carries a factor -1 that the
// epilogue removes. The row order of the A operand is chosen so that the
// output rows of one a-SNP land in one lane of the 32x32 accumulator
// (lane l holds column l%32 and rows (i/4)*8 + (l/32)*4 + i%4), which puts a
// whole pair table in one lane's registers.
//
// FAST tiles (no missing genotype in either SNP range) only need classes {0,1}
// on both sides; the NM row/column of X comes from per-SNP marginals. GENERAL
// tiles carry the NM class on both sides and the q^2 weights for the
// complete-case sum of squares.
#pragma once

#include "epilogue.h"

namespace epi {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int KC = 128;          // individuals per LDS stage
constexpr int ROWB = KC + 16;    // LDS row pitch in bytes (bank-conflict padding)
constexpr int NW = 7;            // weight rows: V, L0, L1, L2, Z0, Z1, Z2

// Identity class map: plane p holds class p (the MFMA form keeps genotype order).
constexpr int CLS_IDENTITY = 0 | (1 << 2) | (2 << 4);

struct Marg {                    // per SNP, per phenotype, over present phenotypes
    int32_t n[3];                // count of class 0, 1, NM
    int32_t cls;                 // plain-form plane -> class map (2 bits per plane)
    long long q[3];              // sum of fixed-point phenotype of class 0, 1, NM
    double ys[3];                // plain form: sum of FP32 centred phenotype of plane 0, 1, NM
    int32_t pn[3];               // plain form: count of plane 0, 1, NM
    int32_t pad;
};

struct MfmaArgs {
    const uint8_t *planes;       // [nsnp][3][Kp] masks
    const int8_t *wts;           // [NW][Kp] for this phenotype
    const Marg *marg;            // [nsnp] for this phenotype
    const int4 *tiles;           // {a0, a1, b0, b1}
    int ntiles;
    int nsnp;
    int Kp;
    int nV;                      // present phenotypes
    long long qV;                // sum q over present
    double sqV;                  // sum q^2 over present
    HitSink hs;
    DumpSink ds;
};

// blockIdx -> tile: consecutive tiles (which share SNP data) go to the same
// XCD. Workgroups are dealt to the 8 XCDs round-robin, so within each full
// group of 256 workgroups, XCD x gets tiles [32x, 32x+32) of the group.
__device__ __forceinline__ int xcd_tile(int bid, int n)
{
    int grp = bid >> 8, r = bid & 255;
    if ((grp << 8) + 256 > n) return bid;
    return (grp << 8) + (r & 7) * 32 + (r >> 3);
}

__device__ __forceinline__ v4i lds_v4(const uint8_t *p) { return *reinterpret_cast<const v4i *>(p); }

// ---------------------------------------------------------------- FAST ----
// Tile 64 a x 64 b, 4 waves; wave w owns a-SNPs [16w, 16w+16) x all 64 b.
// LDS rows: A [plane 0..1][64 a], B [plane 0..1][64 b], W [V, L0, L1, L2].
template <bool DUMP>
__global__ __launch_bounds__(256, 1) void mfma_fast_kernel(const MfmaArgs A)
{
    constexpr int TA = 64, TB = 64, NPL = 2;
    constexpr int ROWS_AB = NPL * (TA + TB);  // 256
    __shared__ __attribute__((aligned(16))) uint8_t lds[(ROWS_AB + 4) * ROWB];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    const int a0 = T.x, a1 = T.y, b0 = T.z, b1 = T.w;
    const int Kp = A.Kp;

    // Global -> LDS copy plan: 2048 16-byte chunks of A/B rows, 8 per thread;
    // 32 chunks of weight rows for threads 0..31.
    const uint8_t *src[8];
    int dst[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int c = tid + 256 * i, row = c >> 3, col = c & 7;
        int side = row >= NPL * TA;                  // 0: A, 1: B
        int r = side ? row - NPL * TA : row;
        int plane = side ? r / TB : r / TA;
        int sl = side ? r % TB : r % TA;
        int snp = (side ? b0 : a0) + sl;
        snp = snp < A.nsnp ? snp : A.nsnp - 1;
        src[i] = A.planes + ((size_t)snp * 3 + plane) * (size_t)Kp + col * 16;
        dst[i] = row * ROWB + col * 16;
    }
    const bool wload = tid < 32;
    const int8_t *wsrc = A.wts + (size_t)(tid >> 3) * Kp + (tid & 7) * 16;
    const int wdst = (ROWS_AB + (tid >> 3)) * ROWB + (tid & 7) * 16;

    // Operand geometry.
    const int m = lane & 31, hf = lane >> 5;
    const int aj = 2 * ((m >> 2) & 1) + (m >> 4);   // a-SNP of this A row within the operand
    const int ag = (m >> 3) & 1;                     // class of this A row
    const int aw = m & 3;                            // weight of this A row

    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (v16i){};

    v4i buf[8], wbuf = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 8; i++) buf[i] = *reinterpret_cast<const v4i *>(src[i]);
    if (wload) wbuf = *reinterpret_cast<const v4i *>(wsrc);

    for (int k0 = 0; k0 < Kp; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; i++) *reinterpret_cast<v4i *>(lds + dst[i]) = buf[i];
        if (wload) *reinterpret_cast<v4i *>(lds + wdst) = wbuf;
        __syncthreads();
        if (k0 + KC < Kp) {
#pragma unroll
            for (int i = 0; i < 8; i++) buf[i] = *reinterpret_cast<const v4i *>(src[i] + k0 + KC);
            if (wload) wbuf = *reinterpret_cast<const v4i *>(wsrc + k0 + KC);
        }
#pragma unroll
        for (int t = 0; t < KC / 32; t++) {
            const int koff = 32 * t + 16 * hf;
            const v4i wr = lds_v4(lds + (ROWS_AB + aw) * ROWB + koff);
            v4i B[4];
#pragma unroll
            for (int bg = 0; bg < 2; bg++)
#pragma unroll
                for (int h = 0; h < 2; h++)
                    B[2 * bg + h] = lds_v4(lds + (NPL * TA + h * TB + 32 * bg + m) * ROWB + koff);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int al = 16 * wid + 4 * i + aj;
                const v4i Aop = lds_v4(lds + (ag * TA + al) * ROWB + koff) & wr;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Aop, B[j], acc[i][j], 0, 0, 0);
            }
        }
    }

    // Per-SNP class marginals of the tile, staged in LDS (the operand buffers
    // are free once every wave has left the K loop).
    __syncthreads();
    int *mg_n = reinterpret_cast<int *>(lds);                    // [128][2]
    long long *mg_q = reinterpret_cast<long long *>(lds + 1024); // [128][2]
    if (tid < TA + TB) {
        int sn = tid < TA ? a0 + tid : b0 + tid - TA;
        sn = sn < A.nsnp ? sn : A.nsnp - 1;
        const Marg &mm = A.marg[sn];
        mg_n[2 * tid] = mm.n[0];
        mg_n[2 * tid + 1] = mm.n[1];
        mg_q[2 * tid] = mm.q[0];
        mg_q[2 * tid + 1] = mm.q[1];
    }
    __syncthreads();

    // Epilogue: lane holds pairs (a = a0 + 16w + 4i + 2hf + p, b = b0 + 32bg + m).
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const int al = 16 * wid + 4 * i + 2 * hf + p;
            const int a = a0 + al;
#pragma unroll
            for (int bg = 0; bg < 2; bg++) {
                const int bl = TA + 32 * bg + m;
                const int b = b0 + 32 * bg + m;
                int xc[3][3];
                long long xs[3][3];
#pragma unroll
                for (int g = 0; g < 2; g++) {
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const v16i &v = acc[i][2 * bg + h];
                        const int e = 8 * p + 4 * g;
                        xc[g][h] = -v[e];
                        xs[g][h] = -((long long)v[e + 1] + 256ll * v[e + 2] + 65536ll * v[e + 3]);
                    }
                    xc[g][2] = mg_n[2 * al + g];
                    xs[g][2] = mg_q[2 * al + g];
                    xc[2][g] = mg_n[2 * bl + g];
                    xs[2][g] = mg_q[2 * bl + g];
                }
                xc[2][2] = A.nV;
                xs[2][2] = A.qV;
                int n[9];
                long long s[9];
                cells_from_x(xc, n);
                cells_from_x(xs, s);
                const bool valid = a < a1 && b < b1 && a > b;
                finish_pair<DUMP>(A.hs, A.ds, valid, a, b, n, s, A.sqV, CLS_IDENTITY, CLS_IDENTITY);
            }
        }
    }
}

// ------------------------------------------------------------- GENERAL ----
// Tile 16 a x 64 b, 4 waves; wave w owns a-SNPs [4w, 4w+4) x all 64 b.
// A operand: 2 a-SNPs x 16 rows: (g in {0,1,NM}) x (V, L0, L1, L2), then
// (NM) x (Z0, Z1, Z2), then a zero row. Lane half hf holds a-SNP hf.
template <bool DUMP>
__global__ __launch_bounds__(256, 1) void mfma_general_kernel(const MfmaArgs A)
{
    constexpr int TA = 16, TB = 64, NPL = 3;
    constexpr int ROWS_AB = NPL * (TA + TB);  // 240
    constexpr int NCHUNK = (ROWS_AB + NW) * (KC / 16);  // 1976
    __shared__ __attribute__((aligned(16))) uint8_t lds[(ROWS_AB + NW + 1) * ROWB];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    const int a0 = T.x, a1 = T.y, b0 = T.z, b1 = T.w;
    const int Kp = A.Kp;

    const uint8_t *src[8];
    int dst[8];
    bool act[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int c = tid + 256 * i, row = c >> 3, col = c & 7;
        act[i] = c < NCHUNK;
        if (row < ROWS_AB) {
            int side = row >= NPL * TA;
            int r = side ? row - NPL * TA : row;
            int plane = side ? r / TB : r / TA;
            int sl = side ? r % TB : r % TA;
            int snp = (side ? b0 : a0) + sl;
            snp = snp < A.nsnp ? snp : A.nsnp - 1;
            src[i] = A.planes + ((size_t)snp * 3 + plane) * (size_t)Kp + col * 16;
        } else {
            int wr = row - ROWS_AB;
            wr = wr < NW ? wr : NW - 1;
            src[i] = reinterpret_cast<const uint8_t *>(A.wts) + (size_t)wr * Kp + col * 16;
        }
        dst[i] = row * ROWB + col * 16;
    }
    // zero row used by the padding A row
    if (tid < KC / 16) *reinterpret_cast<uint4 *>(lds + (ROWS_AB + NW) * ROWB + tid * 16) = make_uint4(0, 0, 0, 0);

    const int m = lane & 31, hf = lane >> 5;
    const int aj = (m >> 2) & 1;          // a-SNP within the operand
    const int q = m >> 3, r4 = m & 3;     // row group and item
    const int aplane = q < 3 ? q : 2;
    const int wrow = q < 3 ? r4 : (r4 < 3 ? 4 + r4 : NW);  // NW = zero row

    v16i acc[2][6];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) acc[i][j] = (v16i){};

    v4i buf[8];
#pragma unroll
    for (int i = 0; i < 8; i++) buf[i] = act[i] ? *reinterpret_cast<const v4i *>(src[i]) : (v4i){0, 0, 0, 0};

    for (int k0 = 0; k0 < Kp; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (act[i]) *reinterpret_cast<v4i *>(lds + dst[i]) = buf[i];
        __syncthreads();
        if (k0 + KC < Kp) {
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (act[i]) buf[i] = *reinterpret_cast<const v4i *>(src[i] + k0 + KC);
        }
#pragma unroll
        for (int t = 0; t < KC / 32; t++) {
            const int koff = 32 * t + 16 * hf;
            const v4i wr = lds_v4(lds + (ROWS_AB + wrow) * ROWB + koff);
            v4i B[6];
#pragma unroll
            for (int bg = 0; bg < 2; bg++)
#pragma unroll
                for (int h = 0; h < 3; h++)
                    B[3 * bg + h] = lds_v4(lds + (NPL * TA + h * TB + 32 * bg + m) * ROWB + koff);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int al = 4 * wid + 2 * i + aj;
                const v4i Aop = lds_v4(lds + (aplane * TA + al) * ROWB + koff) & wr;
#pragma unroll
                for (int j = 0; j < 6; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Aop, B[j], acc[i][j], 0, 0, 0);
            }
        }
    }

    // Epilogue: lane holds pairs (a = a0 + 4w + 2i + hf, b = b0 + 32bg + m);
    // acc[i][3bg+h][4g + w] = -X[g][h][w], g = 3 -> Z weights on NM.
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int a = a0 + 4 * wid + 2 * i + hf;
#pragma unroll
        for (int bg = 0; bg < 2; bg++) {
            const int b = b0 + 32 * bg + m;
            int xc[3][3];
            long long xs[3][3];
#pragma unroll
            for (int g = 0; g < 3; g++)
#pragma unroll
                for (int h = 0; h < 3; h++) {
                    const v16i &v = acc[i][3 * bg + h];
                    xc[g][h] = -v[4 * g];
                    xs[g][h] = -((long long)v[4 * g + 1] + 256ll * v[4 * g + 2] + 65536ll * v[4 * g + 3]);
                }
            const v16i &vz = acc[i][3 * bg + 2];
            const long long z = -((long long)vz[12] + 256ll * vz[13] + 65536ll * vz[14]);
            const double sq = (double)z * 8388608.0;
            int n[9];
            long long s[9];
            cells_from_x(xc, n);
            cells_from_x(xs, s);
            const bool valid = a < a1 && b < b1 && a > b;
            finish_pair<DUMP>(A.hs, A.ds, valid, a, b, n, s, sq, CLS_IDENTITY, CLS_IDENTITY);
        }
    }
}

}  // namespace epi
