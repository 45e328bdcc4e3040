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

struct MfmaScan {                // one dataset x one phenotype
    const uint8_t *planes;       // [nsnp][3][Kp] masks
    const int8_t *wts;           // [NW][Kp] for this phenotype
    const Marg *marg;            // [nsnp] for this phenotype
    int nsnp;
    int Kp;
    int nV;                      // present phenotypes
    int tag;                     // written into the hits
    long long qV;                // sum q over present
    double qVd;                  // the same, as a double (exact)
    double sqV;                  // sum q^2 over present
};

struct MfmaArgs {
    MfmaScan S;
    const int4 *tiles;           // {a0, a1, b0, b1}
    int ntiles;
    HitSink hs;
    DumpSink ds;
};

constexpr int FAST_LDS = 2 * (2 * (64 + 64) + 4) * ROWB;     // largest FAST tile, two stage buffers

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
// Tile TA a x 64 b (TA = 64 or 32), 4 waves; wave w owns a-SNPs
// [TA/4 w, TA/4 (w+1)) x all 64 b, as TA/16 blocks of 32 A rows.
// LDS rows per stage: A [plane 0..1][TA a], B [plane 0..1][64 b],
// W [V, L0, L1, L2]. Two stage buffers; operands of the next 32-individual
// step are read from LDS while the MFMAs of the current step run.
// TA = 64 holds 256 accumulator registers per lane (one wave per SIMD);
// TA = 32 holds 128, so two workgroups fit on a CU.
template <int TA, bool DUMP>
__device__ __forceinline__ void mfma_fast_tile(const MfmaScan &A, const HitSink &hs, const DumpSink &dsk, int a0,
                                               int a1, int b0, int b1, uint8_t *lds)
{
    static_assert(TA == 64 || TA == 32, "fast tile height");
    constexpr int TB = 64, NPL = 2;
    constexpr int MB = TA / 16;               // 32-row A blocks per wave
    constexpr int APW = TA / 4;               // a-SNPs per wave
    constexpr int ROWS_AB = NPL * (TA + TB);  // 256 or 192
    constexpr int NCH = ROWS_AB * 8 / 256;    // 16-byte chunks per thread per stage
    constexpr int STAGE = (ROWS_AB + 4) * ROWB;
    static_assert(2 * STAGE <= FAST_LDS, "LDS size");

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int Kp = A.Kp;

    // Global -> LDS copy plan: ROWS_AB x 8 16-byte chunks of A/B rows, NCH per
    // thread; 32 chunks of weight rows for threads 0..31.
    const uint8_t *src[NCH];
    int dst[NCH];
#pragma unroll
    for (int i = 0; i < NCH; i++) {
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
    const int offW = (ROWS_AB + aw) * ROWB + 16 * hf;
    const int offB = (NPL * TA + m) * ROWB + 16 * hf;          // + (h*TB + 32*bg) * ROWB
    const int offA = (ag * TA + APW * wid + aj) * ROWB + 16 * hf; // + 4*i*ROWB

    v16i acc[MB][4];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (v16i){};

    v4i buf[NCH], wbuf = {0, 0, 0, 0};
    auto gload = [&](int k) {
#pragma unroll
        for (int i = 0; i < NCH; i++) buf[i] = *reinterpret_cast<const v4i *>(src[i] + k);
        if (wload) wbuf = *reinterpret_cast<const v4i *>(wsrc + k);
    };
    auto lstore = [&](uint8_t *st) {
#pragma unroll
        for (int i = 0; i < NCH; i++) *reinterpret_cast<v4i *>(st + dst[i]) = buf[i];
        if (wload) *reinterpret_cast<v4i *>(st + wdst) = wbuf;
    };
    struct Ops {
        v4i w, b[4], am[MB];
    };
    auto fetch = [&](const uint8_t *st, int t, Ops &o) {
        const int k = 32 * t;
        o.w = lds_v4(st + offW + k);
#pragma unroll
        for (int bg = 0; bg < 2; bg++)
#pragma unroll
            for (int h = 0; h < 2; h++) o.b[2 * bg + h] = lds_v4(st + offB + (h * TB + 32 * bg) * ROWB + k);
#pragma unroll
        for (int i = 0; i < MB; i++) o.am[i] = lds_v4(st + offA + 4 * i * ROWB + k);
    };
    auto issue = [&](const Ops &o) {
#pragma unroll
        for (int i = 0; i < MB; i++) {
            const v4i Aop = o.am[i] & o.w;
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Aop, o.b[j], acc[i][j], 0, 0, 0);
        }
    };

    Ops o0, o1;
    gload(0);
    lstore(lds);
    if (KC < Kp) gload(KC);
    __syncthreads();
    fetch(lds, 0, o0);
    for (int k0 = 0, s = 0; k0 < Kp; k0 += KC, s ^= 1) {
        const uint8_t *cur = lds + s * STAGE;
        uint8_t *nxt = lds + (s ^ 1) * STAGE;
        const bool more = k0 + KC < Kp;
        fetch(cur, 1, o1);
        issue(o0);
        fetch(cur, 2, o0);
        if (more) {
            lstore(nxt);
            if (k0 + 2 * KC < Kp) gload(k0 + 2 * KC);
        }
        issue(o1);
        fetch(cur, 3, o1);
        issue(o0);
        // the next stage was written by every wave during this one
        __syncthreads();
        if (more) fetch(nxt, 0, o0);
        issue(o1);
    }

    // Per-SNP class marginals of the tile, staged in LDS (the operand buffers
    // are free once every wave has left the K loop).
    __syncthreads();
    int *mg_n = reinterpret_cast<int *>(lds);                     // [128][2]
    double *mg_q = reinterpret_cast<double *>(lds + 1024);        // [128][2]
    float *mg_f = reinterpret_cast<float *>(lds + 3072);          // [128][2]
    if (tid < TA + TB) {
        int sn = tid < TA ? a0 + tid : b0 + tid - TA;
        sn = sn < A.nsnp ? sn : A.nsnp - 1;
        const Marg &mm = A.marg[sn];
#pragma unroll
        for (int g = 0; g < 2; g++) {
            mg_n[2 * tid + g] = mm.n[g];
            mg_q[2 * tid + g] = (double)mm.q[g];
            mg_f[2 * tid + g] = (float)mm.q[g];
        }
    }
    __syncthreads();
    const float sqf = (float)A.sqV, qVf = (float)A.qV;

    // Epilogue: lane holds pairs (a = a0 + APW w + 4i + 2hf + p, b = b0 + 32bg + m);
    // acc[i][2bg+h][8p + 4g + w] = -X[g][h][w].
#pragma unroll
    for (int i = 0; i < MB; i++) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const int al = APW * wid + 4 * i + 2 * hf + p;
            const int a = a0 + al;
#pragma unroll
            for (int bg = 0; bg < 2; bg++) {
                const int bl = TA + 32 * bg + m;
                const int b = b0 + 32 * bg + m;
                int xc[3][3];
                float xf[3][3];
#pragma unroll
                for (int g = 0; g < 2; g++) {
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const v16i &v = acc[i][2 * bg + h];
                        const int e = 8 * p + 4 * g;
                        xc[g][h] = -v[e];
                        xf[g][h] = -((float)v[e + 1] + 256.0f * (float)v[e + 2] + 65536.0f * (float)v[e + 3]);
                    }
                    xc[g][2] = mg_n[2 * al + g];
                    xf[g][2] = mg_f[2 * al + g];
                    xc[2][g] = mg_n[2 * bl + g];
                    xf[2][g] = mg_f[2 * bl + g];
                }
                xc[2][2] = A.nV;
                xf[2][2] = qVf;
                int n[9];
                float sf[9];
                cells_from_x(xc, n);
                cells_from_x(xf, sf);
                // integer-valued doubles below 2^53: exact
                auto exact = [&](double sd[9], double &sq) {
                    double xs[3][3];
#pragma unroll
                    for (int g = 0; g < 2; g++) {
#pragma unroll
                        for (int h = 0; h < 2; h++) {
                            const v16i &v = acc[i][2 * bg + h];
                            const int e = 8 * p + 4 * g;
                            xs[g][h] = -((double)v[e + 1] + 256.0 * (double)v[e + 2] + 65536.0 * (double)v[e + 3]);
                        }
                        xs[g][2] = mg_q[2 * al + g];
                        xs[2][g] = mg_q[2 * bl + g];
                    }
                    xs[2][2] = A.qVd;
                    cells_from_x(xs, sd);
                    sq = A.sqV;
                };
                const bool valid = a < a1 && b < b1 && a > b;
                finish_pair<DUMP>(hs, dsk, valid, a, b, n, sf, sqf, exact, CLS_IDENTITY, CLS_IDENTITY);
            }
        }
    }
}

template <int TA>
constexpr int fast_lds() { return 2 * (2 * (TA + 64) + 4) * ROWB; }

template <int TA, bool DUMP>
__global__ __launch_bounds__(256, 64 / TA) void mfma_fast_kernel(const MfmaArgs A)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[fast_lds<TA>()];
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    HitSink hs = A.hs;
    hs.tag = A.S.tag;
    mfma_fast_tile<TA, DUMP>(A.S, hs, A.ds, T.x, T.y, T.z, T.w, lds);
}

// ------------------------------------------------------------- GENERAL ----
// Tile 16 a x TB b (TB = 32 NBG = 64 or 32), 4 waves; wave w owns a-SNPs
// [4w, 4w+4) x all TB b. A operand: 2 a-SNPs x 16 rows: (g in {0,1,NM}) x
// (V, L0, L1, L2), then (NM) x (Z0, Z1, Z2), then a zero row. Lane half hf
// holds a-SNP hf. NBG = 2 holds 192 accumulator registers per lane (one wave
// per SIMD), NBG = 1 holds 96 (two workgroups per CU).
template <int NBG>
constexpr int general_lds() { return 2 * (3 * (16 + 32 * NBG) + NW + 1) * ROWB; }

template <int NBG, bool DUMP>
__device__ __forceinline__ void mfma_general_tile(const MfmaScan &A, const HitSink &hs, const DumpSink &dsk, int a0,
                                                  int a1, int b0, int b1, uint8_t *lds)
{
    static_assert(NBG == 1 || NBG == 2, "general tile width");
    constexpr int TA = 16, TB = 32 * NBG, NPL = 3;
    constexpr int ROWS_AB = NPL * (TA + TB);            // 240 or 144
    constexpr int NCHUNK = (ROWS_AB + NW) * (KC / 16);  // 1976 or 1208
    constexpr int NCH = (NCHUNK + 255) / 256;           // chunks per thread
    constexpr int STAGE = (ROWS_AB + NW + 1) * ROWB;    // + one zero row
    static_assert(2 * STAGE == general_lds<NBG>(), "LDS size");

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int Kp = A.Kp;

    const uint8_t *src[NCH];
    int dst[NCH];
    bool act[NCH];
#pragma unroll
    for (int i = 0; i < NCH; i++) {
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
    // zero row used by the padding A row, in both stage buffers
    if (tid < 2 * (KC / 16))
        *reinterpret_cast<v4i *>(lds + (tid >> 3) * STAGE + (ROWS_AB + NW) * ROWB + (tid & 7) * 16) =
            (v4i){0, 0, 0, 0};

    const int m = lane & 31, hf = lane >> 5;
    const int aj = (m >> 2) & 1;          // a-SNP within the operand
    const int q = m >> 3, r4 = m & 3;     // row group and item
    const int aplane = q < 3 ? q : 2;
    const int wrow = q < 3 ? r4 : (r4 < 3 ? 4 + r4 : NW);  // NW = zero row
    const int offW = (ROWS_AB + wrow) * ROWB + 16 * hf;
    const int offB = (NPL * TA + m) * ROWB + 16 * hf;                  // + (h*TB + 32*bg) * ROWB
    const int offA = (aplane * TA + 4 * wid + aj) * ROWB + 16 * hf;    // + 2*i*ROWB

    v16i acc[2][3 * NBG];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3 * NBG; j++) acc[i][j] = (v16i){};

    v4i buf[NCH];
    auto gload = [&](int k) {
#pragma unroll
        for (int i = 0; i < NCH; i++)
            if (act[i]) buf[i] = *reinterpret_cast<const v4i *>(src[i] + k);
    };
    auto lstore = [&](uint8_t *st) {
#pragma unroll
        for (int i = 0; i < NCH; i++)
            if (act[i]) *reinterpret_cast<v4i *>(st + dst[i]) = buf[i];
    };
    struct Ops {
        v4i w, b[3 * NBG], am[2];
    };
    auto fetch = [&](const uint8_t *st, int t, Ops &o) {
        const int k = 32 * t;
        o.w = lds_v4(st + offW + k);
#pragma unroll
        for (int bg = 0; bg < NBG; bg++)
#pragma unroll
            for (int h = 0; h < 3; h++) o.b[3 * bg + h] = lds_v4(st + offB + (h * TB + 32 * bg) * ROWB + k);
#pragma unroll
        for (int i = 0; i < 2; i++) o.am[i] = lds_v4(st + offA + 2 * i * ROWB + k);
    };
    auto issue = [&](const Ops &o) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const v4i Aop = o.am[i] & o.w;
#pragma unroll
            for (int j = 0; j < 3 * NBG; j++)
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Aop, o.b[j], acc[i][j], 0, 0, 0);
        }
    };

    Ops o0, o1;
    gload(0);
    lstore(lds);
    if (KC < Kp) gload(KC);
    __syncthreads();
    fetch(lds, 0, o0);
    for (int k0 = 0, s = 0; k0 < Kp; k0 += KC, s ^= 1) {
        const uint8_t *cur = lds + s * STAGE;
        uint8_t *nxt = lds + (s ^ 1) * STAGE;
        const bool more = k0 + KC < Kp;
        fetch(cur, 1, o1);
        issue(o0);
        fetch(cur, 2, o0);
        if (more) {
            lstore(nxt);
            if (k0 + 2 * KC < Kp) gload(k0 + 2 * KC);
        }
        issue(o1);
        fetch(cur, 3, o1);
        issue(o0);
        __syncthreads();
        if (more) fetch(nxt, 0, o0);
        issue(o1);
    }

    // Epilogue: lane holds pairs (a = a0 + 4w + 2i + hf, b = b0 + 32bg + m);
    // acc[i][3bg+h][4g + w] = -X[g][h][w], g = 3 -> Z weights on NM.
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int a = a0 + 4 * wid + 2 * i + hf;
#pragma unroll
        for (int bg = 0; bg < NBG; bg++) {
            const int b = b0 + 32 * bg + m;
            int xc[3][3];
            float xf[3][3];
#pragma unroll
            for (int g = 0; g < 3; g++)
#pragma unroll
                for (int h = 0; h < 3; h++) {
                    const v16i &v = acc[i][3 * bg + h];
                    xc[g][h] = -v[4 * g];
                    xf[g][h] = -((float)v[4 * g + 1] + 256.0f * (float)v[4 * g + 2] + 65536.0f * (float)v[4 * g + 3]);
                }
            const v16i &vz = acc[i][3 * bg + 2];
            const float sqf = -((float)vz[12] + 256.0f * (float)vz[13] + 65536.0f * (float)vz[14]) * 8388608.0f;
            int n[9];
            float sf[9];
            cells_from_x(xc, n);
            cells_from_x(xf, sf);
            // integer-valued doubles below 2^53: exact
            auto exact = [&](double sd[9], double &sq) {
                double xs[3][3];
#pragma unroll
                for (int g = 0; g < 3; g++)
#pragma unroll
                    for (int h = 0; h < 3; h++) {
                        const v16i &v = acc[i][3 * bg + h];
                        xs[g][h] = -((double)v[4 * g + 1] + 256.0 * (double)v[4 * g + 2] + 65536.0 * (double)v[4 * g + 3]);
                    }
                cells_from_x(xs, sd);
                sq = -((double)vz[12] + 256.0 * (double)vz[13] + 65536.0 * (double)vz[14]) * 8388608.0;
            };
            const bool valid = a < a1 && b < b1 && a > b;
            finish_pair<DUMP>(hs, dsk, valid, a, b, n, sf, sqf, exact, CLS_IDENTITY, CLS_IDENTITY);
        }
    }
}

template <int NBG, bool DUMP>
__global__ __launch_bounds__(256, 3 - NBG) void mfma_general_kernel(const MfmaArgs A)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[general_lds<NBG>()];
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    HitSink hs = A.hs;
    hs.tag = A.S.tag;
    mfma_general_tile<NBG, DUMP>(A.S, hs, A.ds, T.x, T.y, T.z, T.w, lds);
}

}  // namespace epi
