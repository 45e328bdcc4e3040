// Persistent batched scan: one launch walks the tiles of many scans (datasets
// of uneven N and M, and several phenotypes per dataset).
//
// The host keeps one work list per tile kind (FAST, GENERAL, SPARSE), each run by its
// own persistent kernel so that a kernel holds the registers of one tile kind
// only. It orders the items of a list so that consecutive items share SNP data
// (tile-major, phenotype-minor) and cuts the list into NQ contiguous queues of
// equal estimated cost, one per XCD. A workgroup pulls items from the queue of
// its XCD (workgroups are dealt to XCDs round-robin, so blockIdx % 8 is its
// XCD), one item ahead so that the atomic overlaps the current tile, and, when
// that queue is empty, from the others in turn. Every pull is one atomicAdd; a
// claimed slot below its queue's end is always processed, and a pull past the
// end of every queue ends the workgroup, so all workgroups exit once the items
// run out. The queue heads are reset by a
// kernel (not a memset) so that the reset and the scan can be captured in one
// hipGraph and replayed.
#pragma once

#include "kern_mfma.h"
#include "kern_plain.h"

namespace epi {

constexpr int NQ = 8;     // work queues, one per XCD
constexpr int NKIND = 3;  // tile kinds: 0 FAST, 1 GENERAL, 2 SPARSE

struct BatchArgs {
    const MfmaScan *mscan;   // [nscan]
    const PlainScan *pscan;  // [nscan]
    const int4 *work;        // {a0, b0, alen | blen << 8 | general << 16, scan}
    int *qpos;               // [NQ] next item of each queue of this kernel's list (reset to qbegin)
    const int *qend;         // [NQ] one past the last item of each queue
    HitSink hs;
};

// Queue heads of every list ([NKIND][NQ]) and the hit counter.
__global__ void batch_reset(int *qpos, const int *qbegin, unsigned long long *nhits)
{
    const int t = threadIdx.x;
    if (t < NKIND * NQ) qpos[t] = qbegin[t];
    if (t == 0) *nhits = 0ull;
}

// Called by one thread: the next item for a workgroup whose home queue is
// `home`, or -1 when every queue is exhausted.
__device__ __forceinline__ int next_item(const BatchArgs &B, int home)
{
    for (int k = 0; k < NQ; k++) {
        const int q = (home + k) % NQ;
        if (__hip_atomic_load(&B.qpos[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= B.qend[q]) continue;
        const int v = atomicAdd(&B.qpos[q], 1);
        if (v < B.qend[q]) return v;
    }
    return -1;
}

// Threads of a workgroup of the batch kernel for tile kind KIND: the tile's own
// (FAST 4 TA, one wave per SIMD; GENERAL and SPARSE 8 TA).
template <int TA, int KIND>
constexpr int batch_threads() { return KIND == 0 ? fast_threads<TA>() : KIND == 1 ? general_threads<TA>() : sparse_threads<TA>(); }
template <int TA, int KIND>
constexpr int batch_waves_per_simd() { return KIND == 0 ? fast_waves_per_simd<TA>() : 64 / TA; }

// TA: FAST and SPARSE tiles are TA x 64 (TA = 32 or 64), GENERAL tiles
// TA x 32. KIND: the list this kernel runs (0 FAST, 1 GENERAL, 2 SPARSE).
template <int TA, int KIND>
__global__ __launch_bounds__((batch_threads<TA, KIND>()), (batch_waves_per_simd<TA, KIND>())) void mfma_batch_kernel(
    const BatchArgs B)
{
    constexpr int GA = TA;
    constexpr int LDS = KIND == 1 ? general_lds<GA>() : KIND == 2 ? sparse_lds<TA>() : fast_lds<TA>();
    __shared__ __attribute__((aligned(16))) uint8_t lds[LDS];
    __shared__ int item;
    __shared__ int4 wsh;
    const int home = blockIdx.x % NQ;
    const DumpSink nodump = {};
    // Thread 0 claims the next slot of the home queue one item ahead, so the
    // atomic's latency overlaps the current tile.
    int pend = 0;
    if (threadIdx.x == 0) pend = atomicAdd(&B.qpos[home], 1);
    for (;;) {
        __syncthreads();  // the previous tile is done with LDS and the shared item
        if (threadIdx.x == 0) {
            const int it = pend < B.qend[home] ? pend : next_item(B, home);
            item = it;
            if (it >= 0) {
                wsh = B.work[it];
                if (pend < B.qend[home]) pend = atomicAdd(&B.qpos[home], 1);
            }
        }
        __syncthreads();
        const int it = item;
        if (it < 0) break;
        const int4 w = wsh;
        const int a0 = __builtin_amdgcn_readfirstlane(w.x), b0 = __builtin_amdgcn_readfirstlane(w.y);
        const int ext = __builtin_amdgcn_readfirstlane(w.z), sc = __builtin_amdgcn_readfirstlane(w.w);
        // the scan's descriptor stays in memory (its fields load where used)
        // rather than held in registers across the tile
        const MfmaScan &S = B.mscan[sc];
        HitSink hs = B.hs;
        hs.tag = sc;
        const int a1 = a0 + (ext & 255), b1 = b0 + ((ext >> 8) & 255);
        if (KIND == 1)
            mfma_general_tile<GA, false>(S, hs, nodump, a0, a1, b0, b1, lds);
        else if (KIND == 2)
            mfma_sparse_tile<TA, false>(S, hs, nodump, a0, a1, b0, b1, lds);
        else
            mfma_fast_tile<TA, false>(S, hs, nodump, a0, a1, b0, b1, lds);
    }
}

// PA: FAST tiles are 16 PA x 64 (PA = 2 or 1); GENERAL tiles are 16 x 64.
template <int PA, int GEN>
__global__ __launch_bounds__(256) void plain_batch_kernel(const BatchArgs B)
{
    __shared__ int item;
    __shared__ int4 wsh;
    const int home = blockIdx.x % NQ;
    const DumpSink nodump = {};
    int pend = 0;
    if (threadIdx.x == 0) pend = atomicAdd(&B.qpos[home], 1);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const int it = pend < B.qend[home] ? pend : next_item(B, home);
            item = it;
            if (it >= 0) {
                wsh = B.work[it];
                if (pend < B.qend[home]) pend = atomicAdd(&B.qpos[home], 1);
            }
        }
        __syncthreads();
        const int it = item;
        if (it < 0) break;
        const int4 w = wsh;
        const int a0 = __builtin_amdgcn_readfirstlane(w.x), b0 = __builtin_amdgcn_readfirstlane(w.y);
        const int ext = __builtin_amdgcn_readfirstlane(w.z), sc = __builtin_amdgcn_readfirstlane(w.w);
        const PlainScan S = B.pscan[sc];
        HitSink hs = B.hs;
        hs.tag = sc;
        const int a1 = a0 + (ext & 255), b1 = b0 + ((ext >> 8) & 255);
        if (GEN)
            plain_tile<1, 4, false, false>(S, hs, nodump, a0, a1, b0, b1);
        else
            plain_tile<PA, 4, true, false>(S, hs, nodump, a0, a1, b0, b1);
    }
}

}  // namespace epi
