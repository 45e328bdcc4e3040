// MFMA contraction form of the pair scan (gfx950, v_mfma_i32_32x32x32_i8).
//
// For a tile of a-SNPs x b-SNPs the cell table is one INT8 contraction over
// individuals k:
//     X[g][h][w](a, b) = sum_k  (I_g(a,k) * W_w(k)) * I_h(b,k)
// where I_g are 0/1 class indicators (g in {0, 1, NM}) and W_w the per-
// individual weights: W_0 = V (phenotype present, 0/1), W_1..3 = the three
// signed 8-bit limbs of the fixed-point phenotype q = L0 + 256 L1 + 65536 L2,
// W_4..7 the four limbs of round(q^2 / 2^zbits). Every product and every INT32
// sum is exact, so counts are bit-exact and cell sums are exact integers in
// units of the fixed-point phenotype.
//
// Outliers: the fixed-point scale is set by the largest encoded value, so a
// few values far out in the tail would coarsen every other individual. Up to
// OUT_MAX such values per phenotype are carried exactly instead: in the
// contraction they have V = 1 (they count) and zero value limbs, and their
// values are added to each pair's table in the epilogue (add_outliers).
//
// Operands: class masks are stored as bitplanes in HBM ([plane][word][snp],
// bit k of word w = individual 32w + k, planes: class 0, class 1, call
// present) and expanded to bytes on their way into LDS, 128 individuals per
// stage: HBM and L2 carry one bit per individual, the MFMA reads bytes. An A
// operand row is (a, g, w): mask(a, g) AND W_w, with A masks expanded to
// 0xFF / 0x00. A B operand is the mask of 32 b-SNPs for one class, expanded
// to 0x01 / 0x00. The row order of the A operand is chosen so that the
// output rows of one a-SNP land in one lane of the 32x32 accumulator
// (lane l holds column l%32 and rows (i/4)*8 + (l/32)*4 + i%4), which puts a
// whole pair table in one lane's registers.
//
// FAST tiles (no missing genotype in either SNP range) only need classes {0,1}
// on both sides; the NM row/column of X comes from per-SNP marginals. GENERAL
// tiles carry the NM class on both sides and the q^2 weights for the
// complete-case sum of squares.
#pragma once

#include <utility>

#include "epilogue.h"

namespace epi {

// f(std::integral_constant<int, k>) for k = 0 .. N-1, as N separate calls:
// every index is a compile-time constant, whatever the unroller decides.
template <class F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int KC = 128;          // individuals per LDS stage
constexpr int ROWB = KC + 16;    // LDS row pitch in bytes (bank-conflict padding)
constexpr int NWP = 8;           // weight rows: V, L0, L1, L2, Z0, Z1, Z2, Z3
constexpr int OUT_MAX = 32;      // outlier individuals carried exactly, per phenotype

// Identity class map: plane p holds class p (the MFMA form keeps genotype order).
constexpr int CLS_IDENTITY = 0 | (1 << 2) | (2 << 4);

struct Marg {                    // per SNP, per phenotype, over present phenotypes
    int32_t n[3];                // count of class 0, 1, NM
    int32_t cls;                 // plain-form plane -> class map (2 bits per plane)
    long long q[3];              // sum of fixed-point phenotype of class 0, 1, NM
    double ys[3];                // plain form: sum of FP32 centred phenotype of plane 0, 1, NM
    int32_t pn[3];               // plain form: count of plane 0, 1, NM
    int32_t pad;
    long long zs;                // sum of z = round(q^2 / 2^zbits) over NM
};

struct MissPhen {                // per individual, for the SPARSE tiles' missing-call sums
    int32_t v;                   // phenotype present (0 or 1)
    float qf;                    // q as FP32
    float zf;                    // z as FP32
    int32_t q;                   // fixed-point phenotype (0 where missing and for outliers)
    int32_t z;                   // round(q^2 / 2^zbits)
    int32_t pad[3];
};

struct OutVals {                 // values of the outlier individuals, in fixed-point units
    double d[OUT_MAX];
    float f[OUT_MAX];
};

struct MfmaScan {                // one dataset x one phenotype
    const uint32_t *bits;        // [3][Kw][nsnp] bitplanes: class 0, class 1, call present
    const int8_t *wts;           // [NWP][Kp] for this phenotype
    const Marg *marg;            // [nsnp] for this phenotype
    const uint32_t *ocodes;      // [ceil(nout/16)][nsnp] 2-bit codes of the outliers (class, 3 = no call)
    const OutVals *ov;           // outlier values
    int nout;                    // outliers (0..OUT_MAX)
    int nsnp;
    int Kp;                      // = 32 Kw: row stride of wts, plane stride of bits (32 Kp / 32 dwords)
    int Kl;                      // individuals the K loop covers: nid rounded up to KC (<= Kp)
    int nV;                      // present phenotypes
    int tag;                     // written into the hits
    long long qV;                // sum q over present (outliers excluded)
    double qVd;                  // the same, as a double (exact)
    double sqV;                  // sum q^2 over present (outliers excluded)
    double zunit;                // 2^zbits: q^2 units per unit of the Z limbs
    double ouSum, ouSq;          // sum of the outliers' values and of their squares
    // SPARSE tiles (null without missing calls):
    const uint64_t *tbits;       // [nw][nid][4]: bit v % 64 of word (v / 64, i, plane) = plane bit of the SNP
                                 //   with virtual index v at individual i (planes: class 0, class 1, call present)
    const int32_t *vidx;         // [nsnp] virtual index: chromosomes start at multiples of 64
    const long long *moff;       // [nsnp + 1] offsets into mlist
    const int32_t *mlist;        // per SNP, ascending: the individuals with a missing call
    const MissPhen *mp;          // [Kp] for this phenotype
    long long tstride;           // uint64 words per 64-SNP column of tbits (4 nid)
};

// Class of outlier 16w+t at a SNP from its code word w (3: no call).
__device__ __forceinline__ uint32_t out_class(uint32_t code, int t) { return (code >> (2 * t)) & 3u; }

// FAST tiles (no missing call in the tile): every outlier is a complete case of
// every pair, so only the four cells of classes {0,1} x {0,1} depend on the
// pair. Adds sum_j u_j [g_a(j) = g][g_b(j) = h] to x[g][h] (FP64 values for
// the decision, FP32 for the screen).
template <class T>
__device__ __forceinline__ void add_outlier_cells(const MfmaScan &S, int a, int b, T x[3][3])
{
    const int an = a < S.nsnp ? a : S.nsnp - 1, bn = b < S.nsnp ? b : S.nsnp - 1;
    for (int w = 0; 16 * w < S.nout; w++) {
        const uint32_t ca = S.ocodes[(size_t)w * S.nsnp + an], cb = S.ocodes[(size_t)w * S.nsnp + bn];
        const int nt = S.nout - 16 * w < 16 ? S.nout - 16 * w : 16;
        for (int t = 0; t < nt; t++) {
            const uint32_t ga = out_class(ca, t), gb = out_class(cb, t);
            T u;
            if constexpr (sizeof(T) == 8)
                u = S.ov->d[16 * w + t];
            else
                u = S.ov->f[16 * w + t];
            const T u0 = ga == 0u ? u : (T)0, u1 = ga == 1u ? u : (T)0;
            x[0][0] += gb == 0u ? u0 : (T)0;
            x[0][1] += gb == 1u ? u0 : (T)0;
            x[1][0] += gb == 0u ? u1 : (T)0;
            x[1][1] += gb == 1u ? u1 : (T)0;
        }
    }
}

// Adds the outlier individuals that are complete cases of pair (a, b) to its
// {0,1,NM}^2 table x of sums and to the sum of squares sq (fixed-point units).
// Word w of a SNP's codes holds outliers 16w..16w+15, two bits each: the
// genotype class, or 3 for a missing call. The loop bound is wave-uniform.
template <class T>
__device__ __forceinline__ void add_outliers(const MfmaScan &S, int a, int b, T x[3][3], T &sq)
{
    const int an = a < S.nsnp ? a : S.nsnp - 1, bn = b < S.nsnp ? b : S.nsnp - 1;
    for (int w = 0; 16 * w < S.nout; w++) {
        const uint32_t ca = S.ocodes[(size_t)w * S.nsnp + an], cb = S.ocodes[(size_t)w * S.nsnp + bn];
        const int nt = S.nout - 16 * w < 16 ? S.nout - 16 * w : 16;
        for (int t = 0; t < nt; t++) {
            const uint32_t ga = out_class(ca, t), gb = out_class(cb, t);
            T v;
            if constexpr (sizeof(T) == 8)
                v = S.ov->d[16 * w + t];
            else
                v = S.ov->f[16 * w + t];
            const T u = (ga != 3u && gb != 3u) ? v : (T)0;
            const T u0 = ga == 0u ? u : (T)0, u1 = ga == 1u ? u : (T)0;
            x[0][0] += gb == 0u ? u0 : (T)0;
            x[0][1] += gb == 1u ? u0 : (T)0;
            x[1][0] += gb == 0u ? u1 : (T)0;
            x[1][1] += gb == 1u ? u1 : (T)0;
            x[0][2] += u0;
            x[1][2] += u1;
            x[2][0] += gb == 0u ? u : (T)0;
            x[2][1] += gb == 1u ? u : (T)0;
            x[2][2] += u;
            sq += u * u;
        }
    }
}

struct MfmaArgs {
    MfmaScan S;
    const int4 *tiles;           // {a0, a1, b0, b1}
    int ntiles;
    HitSink hs;
    DumpSink ds;
};

constexpr int FAST_LDS = 2 * (2 * (64 + 64) + 5) * ROWB;     // largest FAST tile, two stage buffers
constexpr int FAST_MG_LDS = 4096;                             // a FAST tile's marginals (128 SNPs x 32 bytes)
template <int TA>
constexpr int fast_stage_lds() { return 2 * (2 * (TA + 64) + 5) * ROWB; }

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

// 16 bits -> 16 bytes, bit j -> byte j, as 0x01 / 0x00 (B operands) or
// 0xFF / 0x00 (A operands, which are ANDed with weight bytes). For a nibble
// n, n * 0x00204081 = n + (n << 7) + (n << 14) + (n << 21) puts bit i of n
// at bit 8i (the four shifted copies do not overlap), so masking with
// 0x01010101 leaves one 0/1 per byte. The 0xFF form is a byte permute that
// uses those 0/1 bytes as selectors into the constant 0x0000FF00 (byte 0 is
// 0x00, byte 1 is 0xFF; both halves of the permute's 64-bit source hold it):
// one full-rate instruction, where (t << 8) - t is folded by the compiler into
// a quarter-rate 32-bit multiply.
__device__ __forceinline__ uint32_t widen_ff(uint32_t t) { return __builtin_amdgcn_perm(0xFF00u, 0xFF00u, t); }

template <bool FF>
__device__ __forceinline__ v4i expand16(uint32_t x)
{
    v4i r;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const uint32_t t = (((x >> (4 * d)) & 0xFu) * 0x00204081u) & 0x01010101u;
        r[d] = (int)(FF ? widen_ff(t) : t);
    }
    return r;
}

__device__ __forceinline__ v4i expand16_sel(uint32_t x, bool ff)
{
    v4i r;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const uint32_t t = (((x >> (4 * d)) & 0xFu) * 0x00204081u) & 0x01010101u;
        r[d] = (int)(ff ? widen_ff(t) : t);
    }
    return r;
}

// Staging of one tile's operands for a stage of KC individuals: NB dwords of
// bitplanes per thread (A rows first, then B rows; thread-consecutive along
// SNPs so that loads coalesce; each dword, 32 individuals of one SNP and
// plane, becomes 32 bytes of one LDS row), and the NWR weight rows in 8-byte
// pieces by the first NWR / 4 waves. Two register sets hold the stages in
// flight, so that a stage's loads are issued two stages before it is written
// to LDS. A rows are expanded to 0xFF bytes, B rows to 0x01 bytes. Loads and
// stores are branch-free within a wave, so that a stage is straight-line code
// the scheduler can interleave with MFMAs: waves that load no weights write
// their (in-range, unused) weight load to a dummy LDS row that no operand
// fetch reads. When the dwords do not fill the last round of threads, the
// idle lanes form whole waves, which skip that round with a wave-uniform
// branch; the lanes' dummy destination covers any other case.
template <int NPL, int TA, int TB, int NB, int NWR, int NT>
struct Stager {
    static constexpr int NWD = KC / 32;                 // words per row per stage
    static constexpr int NA = NPL * TA * NWD;           // A-row dwords per stage
    static constexpr int NDW = NPL * (TA + TB) * NWD;   // dwords per stage
    static constexpr int ROWS = NPL * (TA + TB);        // weight rows follow the SNP rows
    static constexpr int DUMMY = ROWS + NWR;            // written by idle lanes, never read
    static_assert(NB * NT >= NDW && (NB - 1) * NT < NDW, "dwords per thread");
    static_assert(NWR % 4 == 0 && NWR * KC / 8 == 64 * (NWR / 4), "weight rows per wave");
    static constexpr bool PART = NB * NT > NDW;         // the last dword round leaves lanes idle
    static constexpr int NBD = NB;                      // dwords per thread per stage
    static_assert((NDW - (NB - 1) * NT) % 64 == 0, "idle lanes fill whole waves");
    bool pact;                                          // this wave has a dword in the last round
    const uint32_t *src[NB];
    int dst[NB];
    bool act[NB], isa[NB];
    const int8_t *wsrc;
    int wdst;
    bool wload;
    int nsnp;
    uint32_t v[2][NB];
    int2 w[2];
    __device__ __forceinline__ Stager(const uint32_t *bits, const int8_t *wts, int Kp, int nsnp_, int a0, int b0,
                                      int tid)
    {
        nsnp = nsnp_;
        const int Kw = Kp / 32;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            int c = tid + NT * i;
            act[i] = c < NDW;
            c = c < NDW ? c : NDW - 1;
            isa[i] = c < NA;
            const int cc = isa[i] ? c : c - NA, T = isa[i] ? TA : TB;
            const int sl = cc % T, pw = cc / T;
            const int plane = pw / NWD, wd = pw % NWD;
            int snp = isa[i] ? a0 + sl : b0 + sl;
            snp = snp < nsnp ? snp : nsnp - 1;
            src[i] = bits + ((size_t)plane * Kw + wd) * nsnp + snp;
            const int row = isa[i] ? plane * TA + sl : NPL * TA + plane * TB + sl;
            dst[i] = act[i] ? row * ROWB + wd * 32 : DUMMY * ROWB;
        }
        pact = (__builtin_amdgcn_readfirstlane(tid) & ~63) + (NB - 1) * NT < NDW;
        const int wv = tid >> 6, r = tid >> 4, ch = tid & 15;
        wload = wv < NWR / 4;
        wsrc = wts + (size_t)(wload ? r : 0) * Kp + ch * 8;
        wdst = wload ? (ROWS + r) * ROWB + ch * 8 : DUMMY * ROWB + 32 + 8 * (tid & 7);
        w[0] = w[1] = make_int2(0, 0);
    }
    // stage starting at individual k0 (a multiple of KC) into register set S
    template <int S>
    __device__ __forceinline__ void load(int k0)
    {
        const size_t off = (size_t)(k0 / 32) * nsnp;
#pragma unroll
        for (int i = 0; i < NB; i++)
            if (!PART || i < NB - 1 || pact) v[S][i] = src[i][off];
        w[S] = *reinterpret_cast<const int2 *>(wsrc + k0);
    }
    // half-dwords [H0, H1) of register set S (half h: 16 bits of dword h / 2,
    // the low half first, 16 bytes of its LDS row), and the weight piece when W
    template <int S, int H0, int H1, bool W>
    __device__ __forceinline__ void store_halves(uint8_t *st) const
    {
#pragma unroll
        for (int h = H0; h < H1; h++) {
            const int i = h / 2;
            if (PART && i == NB - 1 && !pact) continue;
            const uint32_t x = h % 2 ? v[S][i] >> 16 : v[S][i] & 0xFFFFu;
            v4i r;
            if (NT * (i + 1) <= NA)
                r = expand16<true>(x);
            else if (NT * i >= NA)
                r = expand16<false>(x);
            else
                r = expand16_sel(x, isa[i]);
            *reinterpret_cast<v4i *>(st + dst[i] + 16 * (h % 2)) = r;
        }
        if (W) *reinterpret_cast<int2 *>(st + wdst) = w[S];
    }
    // dwords [I0, I1) of register set S, and the weight piece when W
    template <int S, int I0 = 0, int I1 = NB, bool W = true>
    __device__ __forceinline__ void store(uint8_t *st) const
    {
#pragma unroll
        for (int i = I0; i < I1; i++) {
            if (PART && i == NB - 1 && !pact) continue;
            v4i lo, hi;
            if (NT * (i + 1) <= NA) {
                lo = expand16<true>(v[S][i] & 0xFFFFu);
                hi = expand16<true>(v[S][i] >> 16);
            } else if (NT * i >= NA) {
                lo = expand16<false>(v[S][i] & 0xFFFFu);
                hi = expand16<false>(v[S][i] >> 16);
            } else {
                lo = expand16_sel(v[S][i] & 0xFFFFu, isa[i]);
                hi = expand16_sel(v[S][i] >> 16, isa[i]);
            }
            *reinterpret_cast<v4i *>(st + dst[i]) = lo;
            *reinterpret_cast<v4i *>(st + dst[i] + 16) = hi;
        }
        if (W) *reinterpret_cast<int2 *>(st + wdst) = w[S];
    }
};

template <int S>
struct SetTag {
    static constexpr int v = S;
};

// The K loop of a tile over Kl individuals (a multiple of KC) in stages
// of KC through two LDS buffers. Within a stage the operands of the next
// 32-individual step are read from LDS (fetch) while the MFMAs of the
// current one run (issue); the next stage is written to the other buffer
// mid-stage and one barrier per stage publishes it. NMF is the number of
// MFMAs per step (per issue).
template <class St, class Ops, int NMF, class Fetch, class Issue>
__device__ __forceinline__ void k_loop(St &sg, uint8_t *lds, int stage_bytes, int Kl, Fetch fetch, Issue issue)
{
    Ops o0, o1 = {};
    sg.template load<0>(0);
    sg.template load<1>(KC < Kl ? KC : 0);
    sg.template store<0>(lds);
    sg.template load<0>(2 * KC < Kl ? 2 * KC : 0);
    __syncthreads();
    // One stage per region between barriers, in four steps. Step t issues
    // the MFMAs of the operands read during step t - 1 (step 0 those of the
    // previous stage's step 3; zero operands before the first stage), reads
    // the next step's operands from LDS, and does a quarter of the wave's
    // staging. Two dwords per thread (64-row tiles): the first and second
    // halves of the next stage's expansion and LDS writes in steps 0 and 1,
    // the global loads of the stage after it in step 2. Three or more (32-row
    // tiles): the expansion and writes by 16-bit halves spread over all four
    // steps, the global loads after the last write. (Each layout measured
    // faster than the other on its tiles: a scan of N = 8000, M = 50000 on
    // 64-row tiles, a batch of 64 scans with N from 300 to 6000 on 32-row
    // tiles.) A scheduling fence closes each step, so that the compiler
    // cannot hoist the next step's ANDs (whose operands are still in flight
    // from LDS) into this step's MFMAs; within a step, each MFMA is followed
    // by up to VS VALU instructions and LDS reads. VS follows the staging load:
    // three for two dwords per thread and stage (64-row tiles), four for three
    // (32-row tiles), the fastest of 2 to 5 for each on the same two workloads.
    // The next step's LDS reads: one after each MFMA for 64-row tiles, two
    // after each of the first half of the MFMAs for 32-row tiles (likewise
    // measured).
    // Masks: 0x008 MFMA, 0x002 VALU, 0x100 LDS read.
    constexpr int H = (St::NBD + 1) / 2;
    constexpr int VS = St::NBD >= 3 ? 4 : 3;
    auto step_end = [&]() {
#pragma unroll
        for (int u = 0; u < NMF; u++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, VS, 0);
            if (St::NBD < 3)
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            else if (u < NMF / 2)
                __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto body = [&](auto tag, const uint8_t *cur, uint8_t *nxt, int kc) {
        constexpr int S = decltype(tag)::v;
        if constexpr (St::NBD >= 3) {
            constexpr int NH = 2 * St::NBD;  // half-dwords per thread
            issue(o1);
            fetch(cur, 0, o0);
            sg.template store_halves<S, 0, NH / 4, false>(nxt);
            step_end();
            issue(o0);
            fetch(cur, 1, o1);
            sg.template store_halves<S, NH / 4, NH / 2, false>(nxt);
            step_end();
            issue(o1);
            fetch(cur, 2, o0);
            sg.template store_halves<S, NH / 2, 3 * NH / 4, false>(nxt);
            step_end();
            issue(o0);
            fetch(cur, 3, o1);
            sg.template store_halves<S, 3 * NH / 4, NH, true>(nxt);
            sg.template load<S>(kc + 3 * KC < Kl ? kc + 3 * KC : Kl - KC);
            step_end();
        } else {
            issue(o1);
            fetch(cur, 0, o0);
            sg.template store<S, 0, H, false>(nxt);
            step_end();
            issue(o0);
            fetch(cur, 1, o1);
            sg.template store<S, H, St::NBD, true>(nxt);
            step_end();
            issue(o1);
            fetch(cur, 2, o0);
            sg.template load<S>(kc + 3 * KC < Kl ? kc + 3 * KC : Kl - KC);
            step_end();
            issue(o0);
            fetch(cur, 3, o1);
            step_end();
        }
        // the next stage was written by every wave during this one
        __syncthreads();
    };
    int k0 = 0;
    for (; k0 + 2 * KC <= Kl; k0 += 2 * KC) {
        body(SetTag<1>{}, lds, lds + stage_bytes, k0);
        body(SetTag<0>{}, lds + stage_bytes, lds, k0 + KC);
    }
    // an odd stage count ends with one stage in buffer 0; what it stages
    // for a next stage lands in buffer 1 and is never read
    if (k0 < Kl) body(SetTag<1>{}, lds, lds + stage_bytes, k0);
    issue(o1);
}

// The same K loop with one operand set: every step issues the MFMAs of the
// operands in registers and then reads the next step's operands into the
// same registers, so that a register is refilled as soon as its last MFMA or
// AND has read it (the issue order decides how early). For tiles whose
// accumulators leave no room for a second operand set.
template <class St, class Ops, int NMF, class Fetch, class Issue>
__device__ __forceinline__ void k_loop1(St &sg, uint8_t *lds, int stage_bytes, int Kl, Fetch fetch, Issue issue)
{
    Ops o = {};
    sg.template load<0>(0);
    sg.template load<1>(KC < Kl ? KC : 0);
    sg.template store<0>(lds);
    sg.template load<0>(2 * KC < Kl ? 2 * KC : 0);
    __syncthreads();
    // As in k_loop, a scheduling fence closes each step. Within a step each
    // MFMA is followed by up to five VALU instructions, and the first half of
    // the MFMAs by two LDS reads each (the next step's operands); the next
    // stage's expansion and LDS writes are split over steps 1 and 2, its
    // global loads follow in step 2. Measured against one scheduling region
    // per stage: GENERAL tiles of a scan of N = 4000, M = 20000 2.0 % faster
    // with one read after each MFMA, 4.5 % with the reads front-loaded; a
    // four-phenotype batch 1.0 and 1.9 % (three or four VALU slots: 2.8 / 2.3 %
    // and 0.3 / 0.0 %, one read per MFMA).
    auto step_end = [&]() {
#pragma unroll
        for (int u = 0; u < NMF; u++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
            if (u < NMF / 2) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto body = [&](auto tag, const uint8_t *cur, uint8_t *nxt, int kc) {
        constexpr int S = decltype(tag)::v;
        issue(o);  // step 3 of the previous stage (zero operands before the first)
        fetch(cur, 0, o);
        step_end();
        issue(o);
        fetch(cur, 1, o);
        sg.template store<S, 0, (St::NBD + 1) / 2, false>(nxt);
        step_end();
        issue(o);
        fetch(cur, 2, o);
        sg.template store<S, (St::NBD + 1) / 2, St::NBD, true>(nxt);
        sg.template load<S>(kc + 3 * KC < Kl ? kc + 3 * KC : Kl - KC);
        step_end();
        issue(o);
        fetch(cur, 3, o);
        step_end();
        // the next stage was written by every wave during this one
        __syncthreads();
    };
    int k0 = 0;
    for (; k0 + 2 * KC <= Kl; k0 += 2 * KC) {
        body(SetTag<1>{}, lds, lds + stage_bytes, k0);
        body(SetTag<0>{}, lds + stage_bytes, lds, k0 + KC);
    }
    // an odd stage count ends with one stage in buffer 0; what it stages
    // for a next stage lands in buffer 1 and is never read
    if (k0 < Kl) body(SetTag<1>{}, lds, lds + stage_bytes, k0);
    issue(o);
}

// ---------------------------------------------------------------- FAST ----
// Tile TA a x 64 b (TA = 32 or 64). Each wave owns APW = 4 MB a-SNPs x all
// 64 b as MB blocks of 32 A rows (4 a-SNPs x 2 classes x 4 weights): 4 MB
// MFMAs per 32-individual step and 64 MB accumulator registers per lane.
// FAST tiles take MB = 4: 16 a-SNPs per wave, TA / 16 waves, 256 accumulators
// (AGPRs) per lane and one wave per SIMD, whose 512 registers leave room for
// two operand sets in flight; 9 LDS operand reads per 16 MFMAs. SPARSE tiles
// take MB = 2 (8 a-SNPs per wave, TA / 8 waves, two waves per SIMD), since
// their missing-call passes need the registers of a second wave's share.
// LDS rows per stage: A [plane 0..1][TA a], B [plane 0..1][64 b],
// W [V, L0, L1, L2], a dummy row. Two stage buffers; operands of the next 32-individual
// step are read from LDS while the MFMAs of the current step run.
template <int TA, int MB>
constexpr int kloop_threads() { return 64 * TA / (4 * MB); }
constexpr int FAST_MB = 2;  // A blocks per wave of a FAST tile
template <int TA>
constexpr int fast_threads() { return kloop_threads<TA, FAST_MB>(); }
// waves per SIMD that a FAST tile's registers allow
template <int TA>
constexpr int fast_waves_per_simd() { return FAST_MB == 4 ? 1 : 64 / TA; }
template <int TA>
constexpr int sparse_threads() { return kloop_threads<TA, 2>(); }

// The K loop of a tile: acc[i][2bg + h][8p + 4g + w] = X[g][h][w] of the
// pair (a0 + 4MB wid + 4i + 2hf + p, b0 + 32bg + m), m = lane & 31, hf = lane >> 5,
// for classes g, h in {0, 1} and weights w = V, L0, L1, L2.
template <int TA, int MB>
__device__ __forceinline__ void fast_kloop(const MfmaScan &A, int a0, int b0, uint8_t *lds, v16i (&acc)[MB][4])
{
    static_assert(TA == 64 || TA == 32, "fast tile height");
    static_assert(MB == 2 || MB == 4, "A blocks per wave");
    constexpr int TB = 64, NPL = 2, NT = kloop_threads<TA, MB>();
    constexpr int APW = 4 * MB;               // a-SNPs per wave
    constexpr int ROWS_AB = NPL * (TA + TB);  // 192 or 256
    constexpr int NCH = ROWS_AB * (KC / 32) / NT;  // bitplane dwords per thread per stage
    constexpr int STAGE = (ROWS_AB + 5) * ROWB;
    static_assert(2 * STAGE <= FAST_LDS, "LDS size");

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int Kp = A.Kp;

    // Global -> LDS: ROWS_AB rows of bitplanes, expanded to bytes, and the 4
    // weight rows V, L0, L1, L2.
    Stager<NPL, TA, TB, NCH, 4, NT> sg(A.bits, A.wts, Kp, A.nsnp, a0, b0, tid);
    static_assert(NCH * NT == Stager<NPL, TA, TB, NCH, 4, NT>::NDW, "whole dwords per thread");

    // Operand geometry.
    const int m = lane & 31, hf = lane >> 5;
    const int aj = 2 * ((m >> 2) & 1) + (m >> 4);   // a-SNP of this A row within the operand
    const int ag = (m >> 3) & 1;                     // class of this A row
    const int aw = m & 3;                            // weight of this A row
    const int offW = (ROWS_AB + aw) * ROWB + 16 * hf;
    const int offB = (NPL * TA + m) * ROWB + 16 * hf;          // + (h*TB + 32*bg) * ROWB
    const int offA = (ag * TA + APW * wid + aj) * ROWB + 16 * hf; // + 4*i*ROWB

#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (v16i){};

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

    k_loop<decltype(sg), Ops, 4 * MB>(sg, lds, STAGE, A.Kl, fetch, issue);
}

template <int TA, bool DUMP>
__device__ __forceinline__ void mfma_fast_tile(const MfmaScan &A, const HitSink &hs, const DumpSink &dsk, int a0,
                                               int a1, int b0, int b1, uint8_t *lds)
{
    constexpr int TB = 64, MB = FAST_MB, APW = 4 * MB;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int m = lane & 31, hf = lane >> 5;
    // Per-SNP class marginals of the tile, staged in their own LDS region
    // before the K loop, whose first barrier publishes them: their loads
    // overlap the K loop's first stage instead of stalling the epilogue.
    uint8_t *mgl = lds + fast_stage_lds<TA>();
    int *mg_n = reinterpret_cast<int *>(mgl);                     // [128][2]
    double *mg_q = reinterpret_cast<double *>(mgl + 1024);        // [128][2]
    float *mg_f = reinterpret_cast<float *>(mgl + 3072);          // [128][2]
    if (tid < TA + TB) {
        int sn = tid < TA ? a0 + tid : b0 + tid - TA;
        sn = sn < A.nsnp ? sn : A.nsnp - 1;
        const Marg &mm = A.marg[sn];
        // the outliers of each class at this SNP (all of them complete cases in a FAST tile)
        double oq[2] = {0.0, 0.0};
        for (int w = 0; 16 * w < A.nout; w++) {
            const uint32_t c = A.ocodes[(size_t)w * A.nsnp + sn];
            const int nt = A.nout - 16 * w < 16 ? A.nout - 16 * w : 16;
            for (int t = 0; t < nt; t++) {
                const uint32_t g = out_class(c, t);
                const double u = A.ov->d[16 * w + t];
                oq[0] += g == 0u ? u : 0.0;
                oq[1] += g == 1u ? u : 0.0;
            }
        }
#pragma unroll
        for (int g = 0; g < 2; g++) {
            mg_n[2 * tid + g] = mm.n[g];
            mg_q[2 * tid + g] = (double)mm.q[g] + oq[g];
            mg_f[2 * tid + g] = (float)((double)mm.q[g] + oq[g]);
        }
    }
    v16i acc[MB][4];
    fast_kloop<TA, MB>(A, a0, b0, lds, acc);
    const float sqV = (float)(A.sqV + A.ouSq), qVf = (float)(A.qVd + A.ouSum);

    // Epilogue: lane holds pairs (a = a0 + APW w + 4i + 2hf + p, b = b0 + 32bg + m);
    // acc[i][2bg+h][8p + 4g + w] = X[g][h][w]. A pair's values are copied out
    // of the accumulators with constant indices (FastX), so that the per-pair
    // path, however the compiler inlines it, never indexes the accumulators.
    struct FastX {
        int n[2][2], l[2][2][3];  // counts and limb sums of the cells {0,1} x {0,1}
        int al, bl;               // the pair's a and b rows in the tile marginals
    };
    auto pair_x = [&](int i, int p, int bg) {
        FastX x;
#pragma unroll
        for (int g = 0; g < 2; g++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const v16i &v = acc[i][2 * bg + h];
                const int e = 8 * p + 4 * g;
                x.n[g][h] = v[e];
                x.l[g][h][0] = v[e + 1];
                x.l[g][h][1] = v[e + 2];
                x.l[g][h][2] = v[e + 3];
            }
        x.al = APW * wid + 4 * i + 2 * hf + p;
        x.bl = TA + 32 * bg + m;
        return x;
    };
    // The outlier count, read once: the scan descriptor may sit in global
    // memory (the batch kernel), where each slot's possible hit append would
    // otherwise make the compiler reload it.
    const int nout = __builtin_amdgcn_readfirstlane(A.nout);
    // FP32 sums of the pair's 3 x 3 table: the limbs, the outliers of the
    // cells {0,1} x {0,1} (all of them complete cases in a FAST tile) and the
    // marginals.
    auto sums_f = [&](const FastX &x, float xf[3][3]) {
#pragma unroll
        for (int g = 0; g < 2; g++)
#pragma unroll
            for (int h = 0; h < 2; h++)
                xf[g][h] = (float)x.l[g][h][0] + 256.0f * (float)x.l[g][h][1] + 65536.0f * (float)x.l[g][h][2];
        if (nout) add_outlier_cells(A, a0 + x.al, b0 + x.bl - TA, xf);
#pragma unroll
        for (int g = 0; g < 2; g++) {
            xf[g][2] = mg_f[2 * x.al + g];
            xf[2][g] = mg_f[2 * x.bl + g];
        }
        xf[2][2] = qVf;
    };
    // Pair x through the screen, the FP64 decision and the hit list (or the
    // dump); keep = false drops it. Every lane calls.
    auto finish = [&](const FastX &x, bool keep) __attribute__((always_inline)) {
        const int al = x.al, bl = x.bl;
        const int a = a0 + al, b = b0 + bl - TA;
        int xc[3][3];
        float xf[3][3];
#pragma unroll
        for (int g = 0; g < 2; g++) {
#pragma unroll
            for (int h = 0; h < 2; h++) xc[g][h] = x.n[g][h];
            xc[g][2] = mg_n[2 * al + g];
            xc[2][g] = mg_n[2 * bl + g];
        }
        xc[2][2] = A.nV;
        sums_f(x, xf);
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
                for (int h = 0; h < 2; h++)
                    xs[g][h] = (double)x.l[g][h][0] + 256.0 * (double)x.l[g][h][1] + 65536.0 * (double)x.l[g][h][2];
                xs[g][2] = mg_q[2 * al + g];
                xs[2][g] = mg_q[2 * bl + g];
            }
            xs[2][2] = A.qVd + A.ouSum;
            sq = A.sqV + A.ouSq;
            if (A.nout) add_outlier_cells(A, a, b, xs);
            cells_from_x(xs, sd);
        };
        const bool valid = keep && a < a1 && b < b1 && a > b;
        finish_pair<DUMP>(hs, dsk, valid, a, b, n, sf, sqV, exact, CLS_IDENTITY, CLS_IDENTITY);
    };
    // pair slot k = (i, p, bg) = (k / 4, k / 2 % 2, k % 2)
    if constexpr (DUMP) {
        static_for<4 * MB>([&](auto k) {
            constexpr int K = decltype(k)::value;
            finish(pair_x(K / 4, K / 2 % 2, K % 2), true);
        });
    } else {
        // Each pair through the FAST first stage of the screen (the scan's N, T
        // and SST are the tile's); only when some lane of the wave keeps its
        // pair does the wave take the per-pair path (full screen, FP64
        // decision, hit append) for it.
        const float Nf = (float)A.nV, T2N = qVf * qVf / Nf;
        // 64-row tiles: a wave whose pairs all have nine occupied cells takes
        // screen_fast9, whose constants are the scan's; any other wave takes
        // screen_fast_maybe. 32-row tiles (which the autotuner picks for
        // batches of small scans, where empty cells are common) take
        // screen_fast_lane for every pair. Each choice measured faster on its
        // tiles (a scan of N = 8000, M = 50000; a batch of 64 scans with N
        // from 300 to 6000).
        const float d2_9 = Nf - 9.0f;
        const bool all9_ok = d2_9 >= 1.0f;
        const float c1 = 8.0f * hs.t1s / d2_9, c2 = hs.test != 'f' ? 4.0f * hs.t2s / d2_9 : -1.0f;
        const float SST = sqV - T2N;
        static_for<4 * MB>([&](auto k) {
                    constexpr int K = decltype(k)::value, bg = K % 2;
                    const FastX x = pair_x(K / 4, K / 2 % 2, bg);
                    const int a = a0 + x.al, b = b0 + 32 * bg + m;
                    float xn[3][3], xf[3][3];
#pragma unroll
                    for (int g = 0; g < 2; g++) {
#pragma unroll
                        for (int h = 0; h < 2; h++) xn[g][h] = (float)x.n[g][h];
                        xn[g][2] = (float)mg_n[2 * x.al + g];
                        xn[2][g] = (float)mg_n[2 * x.bl + g];
                    }
                    xn[2][2] = Nf;
                    sums_f(x, xf);
                    float nf[9], sf[9];
                    cells_from_x(xn, nf);
                    cells_from_x(xf, sf);
                    const bool valid = (a < a1) & (b < b1) & (a > b);
                    bool keep;
                    if constexpr (TA == 32) {
                        keep = valid & screen_fast_lane(nf, sf, Nf, T2N, sqV, hs.test, hs.safe, hs.t1s, hs.t2s);
                    } else {
                        float nmin = nf[0];
#pragma unroll
                        for (int c = 1; c < 9; c++) nmin = fminf(nmin, nf[c]);
                        if (all9_ok && __ballot(!(nmin > 0.0f)) == 0ull)
                            keep = valid & screen_fast9(nf, sf, T2N, SST, sqV, c1, c2);
                        else
                            keep = valid & screen_fast_maybe(nf, sf, Nf, T2N, sqV, hs.test, hs.safe, hs.t1s, hs.t2s);
                    }
                    if (__ballot(keep) != 0ull) finish(x, keep);
                    // keep the scheduler from interleaving the temporaries of several pairs
                    __builtin_amdgcn_sched_barrier(0);
        });
    }
}

template <int TA>
constexpr int fast_lds() { return fast_stage_lds<TA>() + FAST_MG_LDS; }

template <int TA, bool DUMP>
__global__ __launch_bounds__(fast_threads<TA>(), fast_waves_per_simd<TA>()) void mfma_fast_kernel(const MfmaArgs A)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[fast_lds<TA>()];
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    HitSink hs = A.hs;
    hs.tag = A.S.tag;
    mfma_fast_tile<TA, DUMP>(A.S, hs, A.ds, T.x, T.y, T.z, T.w, lds);
}

// ------------------------------------------------------------- SPARSE ----
// Tiles with missing calls in which every SNP misses few calls. The K loop is
// the FAST tile's, over classes {0, 1} x {0, 1}; the NM row and column of X
// then come from per-SNP marginals (M_x = sums over all calls of class x)
// less sums over the few individuals that miss a call in the other SNP of
// the pair (W = V, q):
//   X[g][NM](a, b) = M_g(a) - sum_{i in miss(b)} I_g(a, i) W(i)    g = 0, 1, NM
//   X[NM][h](a, b) = M_h(b) - sum_{i in miss(a)} I_h(b, i) W(i)    h = 0, 1
//   sum of z over complete cases = Z_NM(a) - sum_{i in miss(b)} I_NM(a, i) z(i)
// The sums run on the VALU in two passes with wave-uniform loops: the lanes
// are 64 SNPs of one side of the tile and the loop runs over the miss list of
// one SNP of the other side. The transposed bitplanes give the class bits of
// those 64 SNPs at the individual as one 64-bit word, and each bit is its
// lane's EXEC bit for the adds. Counts are exact. The FP32 sums feed only the
// screen; a pair that passes it recomputes its sums exactly (INT64, one lane
// per pair, sparse_exact) before the FP64 decision.
//   Pass 1 (lanes = the 64 b-SNPs): wave w, for each of its eight a-SNPs,
//   over miss(a): three values per pair into LDS, [wave][a][value][b].
//   Pass 2 (lanes = the TA a-SNPs, lanes >= TA idle): wave w, for b-SNPs w,
//   w + NW, ..., over miss(b): six values per pair into LDS, [value][a][b]
//   with a row pitch of 65 dwords (a pass-2 store puts its lanes 65 dwords
//   apart, which spreads them over all banks).
constexpr int SP_PITCH = 65;    // dwords per row of pass-2 values

// LDS of a SPARSE tile after its K loop: marginals of the TA + 64 SNPs
// (n[3] int, q[3] and z double, q[3] and z float: 60 bytes each), pass-2
// values, pass-1 values.
template <int TA>
constexpr int sp_mg_bytes() { return 60 * (TA + 64); }
template <int TA>
constexpr int sp_r2_bytes() { return 6 * TA * SP_PITCH * 4; }
template <int TA>
constexpr int sp_r1_bytes() { return (TA / 8) * 8 * 3 * 64 * 4; }

template <int TA>
constexpr int sparse_lds()
{
    return 2 * (2 * (TA + 64) + 5) * ROWB > sp_mg_bytes<TA>() + sp_r2_bytes<TA>() + sp_r1_bytes<TA>()
               ? 2 * (2 * (TA + 64) + 5) * ROWB
               : sp_mg_bytes<TA>() + sp_r2_bytes<TA>() + sp_r1_bytes<TA>();
}

// One individual of a pass-2 loop: the lanes whose bit is set in m0, m1, mn
// (class 0, class 1, call present at their SNP) add v, qf (and zf for mn).
// The wave-uniform masks are the lanes' conditions (selects with a scalar
// mask operand), so the adds need no EXEC changes and do not wait on each
// other.
__device__ __forceinline__ void sp_add3(uint64_t m0, uint64_t m1, uint64_t mn, int v, float qf, float zf,
                                        uint32_t &c0, float &s0, uint32_t &c1, float &s1, uint32_t &cn, float &sn,
                                        float &zn)
{
    c0 += __builtin_amdgcn_inverse_ballot_w64(v ? m0 : 0ull);
    c1 += __builtin_amdgcn_inverse_ballot_w64(v ? m1 : 0ull);
    cn += __builtin_amdgcn_inverse_ballot_w64(v ? mn : 0ull);
    s0 += __builtin_amdgcn_inverse_ballot_w64(m0) ? qf : 0.0f;
    s1 += __builtin_amdgcn_inverse_ballot_w64(m1) ? qf : 0.0f;
    sn += __builtin_amdgcn_inverse_ballot_w64(mn) ? qf : 0.0f;
    zn += __builtin_amdgcn_inverse_ballot_w64(mn) ? zf : 0.0f;
}

// One individual of a pass-1 loop (classes 0 and 1 only).
__device__ __forceinline__ void sp_add2(uint64_t m0, uint64_t m1, int v, float qf, uint32_t &c0, float &s0,
                                        uint32_t &c1, float &s1)
{
    c0 += __builtin_amdgcn_inverse_ballot_w64(v ? m0 : 0ull);
    c1 += __builtin_amdgcn_inverse_ballot_w64(v ? m1 : 0ull);
    s0 += __builtin_amdgcn_inverse_ballot_w64(m0) ? qf : 0.0f;
    s1 += __builtin_amdgcn_inverse_ballot_w64(m1) ? qf : 0.0f;
}

// The individuals of a pass are consumed one at a time by every lane (the
// lanes are SNPs), but loaded 64 at a time, one per lane: a chunk is up to 64
// consecutive entries of one miss list with, per lane, its individual's words
// in the tile's 64-SNP column tb (32 bytes per individual: class 0, class 1,
// call present) and its MissPhen head {v, qf, zf}. v_readlane then hands each
// entry's words and values to the scalar unit for the EXEC-masked adds. A
// chunk's loads are issued one chunk ahead, so their latency hides behind the
// previous chunk's adds.
struct SpChunk {
    uint64_t m0, m1, mn;
    int v;
    float qf, zf;
    int n;  // entries in the chunk (wave-uniform)
};

template <int NPL>
__device__ __forceinline__ SpChunk sp_fetch(const MfmaScan &S, long long c, long long e1, const uint8_t *tb)
{
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    typedef int v3 __attribute__((ext_vector_type(3)));
    const int lane = threadIdx.x & 63;
    SpChunk k;
    k.n = (int)(e1 - c < 64 ? e1 - c : 64);
    const uint32_t i = (uint32_t)S.mlist[c + (lane < k.n ? lane : 0)];
    const uint32_t o = i << 5;
    // exactly the words used (v, qf, zf), so that no later write to an unused
    // register of the load has to wait for it
    const u64x2 t = *reinterpret_cast<const u64x2 *>(tb + o);
    const v3 w = *reinterpret_cast<const v3 *>(reinterpret_cast<const uint8_t *>(S.mp) + o);
    k.m0 = t[0];
    k.m1 = t[1];
    k.mn = NPL == 3 ? *reinterpret_cast<const uint64_t *>(tb + o + 16) : 0ull;
    k.v = w[0];
    k.qf = __int_as_float(w[1]);
    k.zf = __int_as_float(w[2]);
    return k;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int u)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), u);
    return ((uint64_t)hi << 32) | lo;
}

// Wave-uniform loop over the entries of nlist miss lists, list t being
// mlist[moff[snp(t)], moff[snp(t) + 1]) (snp(t) < 0: an empty list);
// begin(t) before a list's entries, fn(m0, m1, mn, v, qf, zf) per entry and
// end(t) after them.
template <int NPL, class Snp, class Begin, class Fn, class End>
__device__ __forceinline__ void sp_lists(const MfmaScan &S, int nlist, const uint8_t *tb, Snp snp, Begin begin, Fn fn,
                                         End end)
{
    int t = 0;
    long long c = 0, e1 = 0;
    auto bounds = [&](int tt, long long &cc, long long &ee) {
        const int s = tt < nlist ? snp(tt) : -1;
        cc = s >= 0 ? S.moff[s] : 0;
        ee = s >= 0 ? S.moff[s + 1] : 0;
    };
    // first non-empty chunk
    bounds(0, c, e1);
    SpChunk cur = {};
    if (c < e1) cur = sp_fetch<NPL>(S, c, e1, tb);
    begin(0);
    for (;;) {
        // This chunk's words are in registers before the next chunk's loads go
        // out (the load counter is in order, but a wait placed in the loop
        // below would also wait for the next chunk).
        asm volatile("" ::"v"(cur.m0), "v"(cur.m1), "v"(cur.mn), "v"(cur.v), "v"(cur.qf), "v"(cur.zf) : "memory");
        // the next chunk: later in this list, or the first of a following non-empty list
        int tn = t;
        long long cn = c + 64, en = e1;
        int skipped = 0;
        while (cn >= en && tn < nlist) {
            tn++;
            skipped++;
            bounds(tn, cn, en);
        }
        SpChunk nxt = {};
        if (tn < nlist && cn < en) nxt = sp_fetch<NPL>(S, cn, en, tb);
        if (c < e1)
            for (int u = 0; u < cur.n; u++)
                fn(readlane64(cur.m0, u), readlane64(cur.m1, u), NPL == 3 ? readlane64(cur.mn, u) : 0ull,
                   __builtin_amdgcn_readlane(cur.v, u), __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.qf), u)),
                   NPL == 3 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.zf), u)) : 0.0f);
        // lists finished before the next chunk
        for (int k = 0; k < skipped; k++) {
            end(t + k);
            if (t + k + 1 < nlist) begin(t + k + 1);
        }
        if (tn >= nlist) break;
        t = tn;
        c = cn;
        e1 = en;
        cur = nxt;
    }
}

// Exact (INT64) sums of pair (a, b) over the individuals missing a call in
// one SNP, by the other SNP's class: over miss(b), a's classes 0, 1, NM into
// dq[0..2] (and z over NM into dz); over miss(a), b's classes 0, 1 into eq.
// One lane per pair; the loops are per lane.
__device__ __forceinline__ void sparse_exact(const MfmaScan &S, int a, int b, long long dq[3], long long &dz,
                                             long long eq[2])
{
    const int an = a < S.nsnp ? a : S.nsnp - 1, bn = b < S.nsnp ? b : S.nsnp - 1;
    const size_t ps = (size_t)(S.Kp / 32) * S.nsnp;  // bitplane stride
    dq[0] = dq[1] = dq[2] = dz = eq[0] = eq[1] = 0;
    for (long long e = S.moff[bn]; e < S.moff[bn + 1]; e++) {
        const int i = S.mlist[e];
        const uint32_t bit = 1u << (i & 31);
        const size_t o = (size_t)(i >> 5) * S.nsnp + an;
        const long long q = S.mp[i].q;
        if (S.bits[o] & bit) dq[0] += q;
        if (S.bits[ps + o] & bit) dq[1] += q;
        if (S.bits[2 * ps + o] & bit) {
            dq[2] += q;
            dz += S.mp[i].z;
        }
    }
    for (long long e = S.moff[an]; e < S.moff[an + 1]; e++) {
        const int i = S.mlist[e];
        const uint32_t bit = 1u << (i & 31);
        const size_t o = (size_t)(i >> 5) * S.nsnp + bn;
        const long long q = S.mp[i].q;
        if (S.bits[o] & bit) eq[0] += q;
        if (S.bits[ps + o] & bit) eq[1] += q;
    }
}

template <int TA, bool DUMP>
__device__ __forceinline__ void mfma_sparse_tile(const MfmaScan &A, const HitSink &hs, const DumpSink &dsk, int a0,
                                                 int a1, int b0, int b1, uint8_t *lds)
{
    constexpr int TB = 64, MB = 2, APW = 8, NW = TA / 8;
    static_assert(sparse_lds<TA>() * (64 / TA) <= 160 * 1024, "LDS size");
    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m = lane & 31, hf = lane >> 5;
    v16i acc[MB][4];
    fast_kloop<TA, MB>(A, a0, b0, lds, acc);
    __syncthreads();  // the stage buffers are free

    // Marginals of the tile's SNPs (a-SNPs first), pass-2 and pass-1 values.
    constexpr int NS = TA + TB;
    int *mg_n = reinterpret_cast<int *>(lds);                              // [3][NS]
    double *mg_q = reinterpret_cast<double *>(lds + 12 * NS);              // [3][NS]
    double *mg_z = reinterpret_cast<double *>(lds + 36 * NS);              // [NS]
    float *mg_f = reinterpret_cast<float *>(lds + 44 * NS);                // [3][NS]
    float *mg_zf = reinterpret_cast<float *>(lds + 56 * NS);               // [NS]
    uint32_t *r2u = reinterpret_cast<uint32_t *>(lds + sp_mg_bytes<TA>()); // [6][TA][SP_PITCH]
    float *r2f = reinterpret_cast<float *>(r2u);
    uint32_t *r1u = reinterpret_cast<uint32_t *>(lds + sp_mg_bytes<TA>() + sp_r2_bytes<TA>());  // [NW][8][3][64]
    float *r1f = reinterpret_cast<float *>(r1u);
    if (tid < NS) {
        int sn = tid < TA ? a0 + tid : b0 + tid - TA;
        sn = sn < A.nsnp ? sn : A.nsnp - 1;
        const Marg &mm = A.marg[sn];
#pragma unroll
        for (int g = 0; g < 3; g++) {
            mg_n[NS * g + tid] = mm.n[g];
            mg_q[NS * g + tid] = (double)mm.q[g];
            mg_f[NS * g + tid] = (float)mm.q[g];
        }
        mg_z[tid] = (double)mm.zs;
        mg_zf[tid] = (float)mm.zs;
    }
    // Pass 1: lanes = b-SNPs b0 + lane, over miss(a) for a = a0 + 8 wid + j.
    {
        const int vb = __builtin_amdgcn_readfirstlane(A.vidx[b0]);
        const uint8_t *tb = reinterpret_cast<const uint8_t *>(A.tbits + (size_t)(vb >> 6) * A.tstride);
        uint32_t c0 = 0, c1 = 0;
        float s0 = 0.0f, s1 = 0.0f;
        uint32_t *r1w = r1u + wid * (APW * 3 * 64);
        sp_lists<2>(
            A, APW, tb, [&](int j) { return a0 + APW * wid + j < a1 ? a0 + APW * wid + j : -1; },
            [&](int) {
                c0 = c1 = 0;
                s0 = s1 = 0.0f;
            },
            [&](uint64_t m0, uint64_t m1, uint64_t, int v, float qf, float) { sp_add2(m0, m1, v, qf, c0, s0, c1, s1); },
            [&](int j) {
                r1w[(3 * j + 0) * 64 + lane] = c0 | (c1 << 16);
                reinterpret_cast<float *>(r1w)[(3 * j + 1) * 64 + lane] = s0;
                reinterpret_cast<float *>(r1w)[(3 * j + 2) * 64 + lane] = s1;
            });
    }
    // Pass 2: lanes = a-SNPs a0 + lane, over miss(b) for b = b0 + wid + NW k.
    {
        const int va = __builtin_amdgcn_readfirstlane(A.vidx[a0]);
        const uint8_t *ta = reinterpret_cast<const uint8_t *>(A.tbits + (size_t)(va >> 6) * A.tstride);
        const int ash = va & 63;  // 0, or 32 for the second half of a 64-SNP word (TA = 32)
        uint32_t c0 = 0, c1 = 0, cn = 0;
        float s0 = 0.0f, s1 = 0.0f, sn = 0.0f, zn = 0.0f;
        sp_lists<3>(
            A, TB / NW, ta, [&](int k) { return b0 + wid + NW * k < b1 ? b0 + wid + NW * k : -1; },
            [&](int) {
                c0 = c1 = cn = 0;
                s0 = s1 = sn = zn = 0.0f;
            },
            [&](uint64_t m0, uint64_t m1, uint64_t mn, int v, float qf, float zf) {
                if (TA == 64)
                    sp_add3(m0, m1, mn, v, qf, zf, c0, s0, c1, s1, cn, sn, zn);
                else
                    sp_add3(m0 >> ash, m1 >> ash, mn >> ash, v, qf, zf, c0, s0, c1, s1, cn, sn, zn);
            },
            [&](int k) {
                if (lane < TA) {
                    const int o = lane * SP_PITCH + wid + NW * k;
                    r2u[0 * TA * SP_PITCH + o] = c0 | (c1 << 16);
                    r2u[1 * TA * SP_PITCH + o] = cn;
                    r2f[2 * TA * SP_PITCH + o] = s0;
                    r2f[3 * TA * SP_PITCH + o] = s1;
                    r2f[4 * TA * SP_PITCH + o] = sn;
                    r2f[5 * TA * SP_PITCH + o] = zn;
                }
            });
    }
    __syncthreads();  // marginals and pass values are in LDS

    const float zu = (float)A.zunit;
#pragma unroll
    for (int i = 0; i < MB; i++) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const int al = APW * wid + 4 * i + 2 * hf + p;
            const int a = a0 + al;
#pragma unroll
            for (int bg = 0; bg < 2; bg++) {
                const int bl = 32 * bg + m, tb = TA + bl;
                const int b = b0 + bl;
                const int o = al * SP_PITCH + bl;
                const uint32_t rc = r2u[o], rn = r2u[TA * SP_PITCH + o];
                const float rs0 = r2f[2 * TA * SP_PITCH + o], rs1 = r2f[3 * TA * SP_PITCH + o];
                const float rsn = r2f[4 * TA * SP_PITCH + o], rz = r2f[5 * TA * SP_PITCH + o];
                const int jw = 4 * i + 2 * hf + p;  // a-SNP of the pair among the wave's eight
                const uint32_t *r1p = r1u + (wid * APW + jw) * 3 * 64 + bl;
                const uint32_t dc = r1p[0];
                const float d1s0 = __uint_as_float(r1p[64]), d1s1 = __uint_as_float(r1p[128]);
                int xc[3][3];
                float xf[3][3];
#pragma unroll
                for (int g = 0; g < 2; g++)
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const v16i &v = acc[i][2 * bg + h];
                        const int e = 8 * p + 4 * g;
                        xc[g][h] = v[e];
                        xf[g][h] = (float)v[e + 1] + 256.0f * (float)v[e + 2] + 65536.0f * (float)v[e + 3];
                    }
                xc[0][2] = mg_n[al] - (int)(rc & 0xFFFFu);
                xc[1][2] = mg_n[NS + al] - (int)(rc >> 16);
                xc[2][2] = mg_n[2 * NS + al] - (int)rn;
                xc[2][0] = mg_n[tb] - (int)(dc & 0xFFFFu);
                xc[2][1] = mg_n[NS + tb] - (int)(dc >> 16);
                xf[0][2] = mg_f[al] - rs0;
                xf[1][2] = mg_f[NS + al] - rs1;
                xf[2][2] = mg_f[2 * NS + al] - rsn;
                xf[2][0] = mg_f[tb] - d1s0;
                xf[2][1] = mg_f[NS + tb] - d1s1;
                float sqf = (mg_zf[al] - rz) * zu;
                if (A.nout) add_outliers(A, a, b, xf, sqf);
                int n[9];
                float sf[9];
                cells_from_x(xc, n);
                cells_from_x(xf, sf);
                // integer-valued doubles below 2^53: exact
                auto exact = [&](double sd[9], double &sq) {
                    double xs[3][3];
#pragma unroll
                    for (int g = 0; g < 2; g++)
#pragma unroll
                        for (int h = 0; h < 2; h++) {
                            const v16i &v = acc[i][2 * bg + h];
                            const int e = 8 * p + 4 * g;
                            xs[g][h] = (double)v[e + 1] + 256.0 * (double)v[e + 2] + 65536.0 * (double)v[e + 3];
                        }
                    long long dq[3], dz, eq[2];
                    sparse_exact(A, a, b, dq, dz, eq);
                    xs[0][2] = mg_q[al] - (double)dq[0];
                    xs[1][2] = mg_q[NS + al] - (double)dq[1];
                    xs[2][2] = mg_q[2 * NS + al] - (double)dq[2];
                    xs[2][0] = mg_q[tb] - (double)eq[0];
                    xs[2][1] = mg_q[NS + tb] - (double)eq[1];
                    sq = (mg_z[al] - (double)dz) * A.zunit;
                    if (A.nout) add_outliers(A, a, b, xs, sq);
                    cells_from_x(xs, sd);
                };
                const bool valid = a < a1 && b < b1 && a > b;
                finish_pair<DUMP>(hs, dsk, valid, a, b, n, sf, sqf, exact, CLS_IDENTITY, CLS_IDENTITY);
            }
        }
    }
}

template <int TA, bool DUMP>
__global__ __launch_bounds__(sparse_threads<TA>(), 64 / TA) void mfma_sparse_kernel(const MfmaArgs A)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[sparse_lds<TA>()];
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    HitSink hs = A.hs;
    hs.tag = A.S.tag;
    mfma_sparse_tile<TA, DUMP>(A.S, hs, A.ds, T.x, T.y, T.z, T.w, lds);
}

// ------------------------------------------------------------- GENERAL ----
// Tile GA a x 32 b (GA = 32 or 64) with GA / 8 waves; wave w owns a-SNPs
// [8w, 8w+8) x the 32 b. Three A blocks, one per class g of the a-SNP
// (0, 1, NM), of 32 rows each: row m is a-SNP 4((m>>2)&1) + (m>>3) of the
// wave's eight with weight m & 3 (V, L0, L1, L2), so that lane half hf of the
// 32x32 accumulator holds a-SNPs 4hf..4hf+3 with all four weights. Each A
// block meets the three B classes (0, 1, NM): 9 MFMAs. A fourth block, the
// same NM masks with weights Z0..Z3 (the limbs of q^2 / 2^zbits), meets only
// the NM class of B for the complete-case sum of q^2: 10 MFMAs per
// 32-individual step for 8 x 32 pairs (40 MACs per pair and individual) and
// 160 accumulator registers, two waves per SIMD.
template <int GA>
constexpr int general_lds() { return 2 * (3 * (GA + 32) + NWP + 1) * ROWB; }

template <int GA>
constexpr int general_threads() { return 8 * GA; }

template <int GA, bool DUMP>
__device__ __forceinline__ void mfma_general_tile(const MfmaScan &A, const HitSink &hs, const DumpSink &dsk, int a0,
                                                  int a1, int b0, int b1, uint8_t *lds)
{
    static_assert(GA == 32 || GA == 64, "general tile height");
    constexpr int TA = GA, TB = 32, NPL = 3, NT = general_threads<GA>();
    constexpr int ROWS_AB = NPL * (TA + TB);
    constexpr int NCH = (ROWS_AB * (KC / 32) + NT - 1) / NT;  // bitplane dwords per thread
    constexpr int STAGE = (ROWS_AB + NWP + 1) * ROWB;          // + weight rows, a dummy row
    static_assert(2 * STAGE == general_lds<GA>(), "LDS size");

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int Kp = A.Kp;

    // Global -> LDS: ROWS_AB rows of bitplanes, expanded to bytes, and the NWP
    // weight rows V, L0, L1, L2, Z0, Z1, Z2, Z3.
    Stager<NPL, TA, TB, NCH, NWP, NT> sg(A.bits, A.wts, Kp, A.nsnp, a0, b0, tid);

    const int m = lane & 31, hf = lane >> 5;
    const int aj = 4 * ((m >> 2) & 1) + (m >> 3);  // a-SNP of A row m among the wave's eight
    const int wr = m & 3;                          // weight of A row m
    const int offA = (8 * wid + aj) * ROWB + 16 * hf;                           // + g * TA * ROWB
    const int offW = (ROWS_AB + wr) * ROWB + 16 * hf;                           // V, L0, L1, L2
    const int offZ = (ROWS_AB + 4 + wr) * ROWB + 16 * hf;                       // Z0, Z1, Z2, Z3
    const int offB = (NPL * TA + m) * ROWB + 16 * hf;                           // + h * TB * ROWB

    v16i acc[3][3], accz = (v16i){};
#pragma unroll
    for (int g = 0; g < 3; g++)
#pragma unroll
        for (int h = 0; h < 3; h++) acc[g][h] = (v16i){};

    struct Ops {
        v4i w, z, b[3], am[3];
    };
    auto fetch = [&](const uint8_t *st, int t, Ops &o) {
        const int k = 32 * t;
        o.w = lds_v4(st + offW + k);
        o.z = lds_v4(st + offZ + k);
#pragma unroll
        for (int h = 0; h < 3; h++) o.b[h] = lds_v4(st + offB + h * TB * ROWB + k);
#pragma unroll
        for (int g = 0; g < 3; g++) o.am[g] = lds_v4(st + offA + g * TA * ROWB + k);
    };
    // B class outermost: b[0] and b[1] are free for the next step's reads
    // after the third and sixth MFMA, the A masks and weights after the ANDs.
    auto issue = [&](const Ops &o) {
        v4i Aop[3];
#pragma unroll
        for (int g = 0; g < 3; g++) Aop[g] = o.am[g] & o.w;
        const v4i Az = o.am[2] & o.z;
#pragma unroll
        for (int h = 0; h < 3; h++)
#pragma unroll
            for (int g = 0; g < 3; g++)
                acc[g][h] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Aop[g], o.b[h], acc[g][h], 0, 0, 0);
        accz = __builtin_amdgcn_mfma_i32_32x32x32_i8(Az, o.b[2], accz, 0, 0, 0);
    };

    k_loop1<decltype(sg), Ops, 10>(sg, lds, STAGE, A.Kl, fetch, issue);

    // Epilogue: lane holds pairs (a = a0 + 8w + 4hf + j, b = b0 + m), j = 0..3;
    // acc[g][h][4j + w] = X[g][h][w], accz[4j + w] = X[NM][NM][Z_w].
    const int b = b0 + m;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int a = a0 + 8 * wid + 4 * hf + j;
        int xc[3][3];
        float xf[3][3];
#pragma unroll
        for (int g = 0; g < 3; g++)
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const v16i &v = acc[g][h];
                xc[g][h] = v[4 * j];
                xf[g][h] = (float)v[4 * j + 1] + 256.0f * (float)v[4 * j + 2] + 65536.0f * (float)v[4 * j + 3];
            }
        float sqf = ((float)accz[4 * j] + 256.0f * (float)accz[4 * j + 1] + 65536.0f * (float)accz[4 * j + 2] +
                     16777216.0f * (float)accz[4 * j + 3]) *
                    (float)A.zunit;
        if (A.nout) add_outliers(A, a, b, xf, sqf);
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
                    const v16i &v = acc[g][h];
                    xs[g][h] = (double)v[4 * j + 1] + 256.0 * (double)v[4 * j + 2] + 65536.0 * (double)v[4 * j + 3];
                }
            sq = ((double)accz[4 * j] + 256.0 * (double)accz[4 * j + 1] + 65536.0 * (double)accz[4 * j + 2] +
                  16777216.0 * (double)accz[4 * j + 3]) *
                 A.zunit;
            if (A.nout) add_outliers(A, a, b, xs, sq);
            cells_from_x(xs, sd);
        };
        const bool valid = a < a1 && b < b1 && a > b;
        finish_pair<DUMP>(hs, dsk, valid, a, b, n, sf, sqf, exact, CLS_IDENTITY, CLS_IDENTITY);
    }
}

template <int GA, bool DUMP>
__global__ __launch_bounds__(general_threads<GA>(), 64 / GA) void mfma_general_kernel(const MfmaArgs A)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[general_lds<GA>()];
    const int4 T = A.tiles[xcd_tile(blockIdx.x, A.ntiles)];
    HitSink hs = A.hs;
    hs.tag = A.S.tag;
    mfma_general_tile<GA, DUMP>(A.S, hs, A.ds, T.x, T.y, T.z, T.w, lds);
}

}  // namespace epi
