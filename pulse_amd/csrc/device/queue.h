// queue.h -- pieces shared by the single-view and the batched stage kernels:
// wave-level queue append, the traversal variant switch and the per-frame
// arguments of the path loop.
#pragma once

#include "pt_device.h"
#include "traverse.h"
#include "wide.h"

namespace pt {

constexpr int kMaxTraces = 64;

// A persistent trace launch splits each queue into kQueueParts contiguous
// parts, one per XCD: the block that runs on XCD x (blocks are dispatched
// to the 8 XCDs in turn) takes rays from part x first, so each XCD's L2
// serves the rays of one stretch of the queue (one region of the screen)
// instead of all eight caching the same one; a wave whose part is drained
// goes on to the next part.  Each part's taken-ray head has its own 64-byte
// line, kHeadWords words per queue kind (extension, shadow).
constexpr int kQueueParts = 8;
constexpr int kHeadStride = 16;
constexpr int kHeadWords = (kMaxTraces + 1) * kQueueParts * kHeadStride;

// Traversal variants: 0 = the reference's nested TLAS/BLAS loops,
// 1 = single loop over both levels (traverse.h), same results; 2 = the
// 4-wide BVHs (wide.h), same results up to the order among exactly tied hits.
template <int LOOP, bool ANY_HIT>
__device__ __forceinline__ bool trace_any(const DevScene &sc, V3 o, V3 d, float &t, uint32_t &inst, uint32_t &tri,
                                          float &u, float &v, float threshold, LdsStack &stk) {
    if (LOOP == 2) return trace_wide<ANY_HIT>(sc, o, d, t, inst, tri, u, v, threshold, stk);
    if (LOOP == 1) return trace_unified<ANY_HIT>(sc, o, d, t, inst, tri, u, v, threshold, stk);
    return traverse_tlas<ANY_HIT>(sc, o, d, t, inst, tri, u, v, threshold, stk);
}

struct FrameArgs {
    pulse_camera cam;
    uint32_t mode;
    uint32_t nee_ggx;
    uint32_t max_traces;      // extension traces per path
    uint32_t rr_start;        // first loop depth with Russian roulette
    uint32_t accumulation_count;
    uint32_t horizon_clamp;   // PULSE_FLAG_HORIZON_CLAMP
    uint32_t light_power;     // PULSE_FLAG_LIGHT_POWER
    uint32_t gi;              // the G-buffer GI pass (every trace index is a loop-body hit)
};

// Appends `valid` items of a wave to a queue: returns this lane's slot.
// All 64 lanes of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(bool valid, uint32_t *counter) {
    uint64_t mask = __ballot(valid);
    uint32_t total = __popcll(mask);
    uint32_t lane = __lane_id();
    uint32_t before = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (total > 0) {
        int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        if (static_cast<int>(lane) == leader) base = atomicAdd(counter, total);
        base = __shfl(base, leader);
    }
    return base + before;
}

// Appends the `valid` items of a block of BLOCK threads (all of which call
// it) to two queues at once with one atomic per queue per block: returns
// this thread's slots in r0 / r1.  `lds`: 2 * (BLOCK / 64 + 1) words.
template <int BLOCK>
__device__ __forceinline__ void block_append2(bool v0, uint32_t *c0, bool v1, uint32_t *c1, uint32_t *lds, uint32_t &r0,
                                              uint32_t &r1) {
    constexpr uint32_t kWaves = BLOCK / 64;
    const uint64_t m0 = __ballot(v0), m1 = __ballot(v1);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    if (lane == 0u) {
        lds[wave] = static_cast<uint32_t>(__popcll(m0));
        lds[kWaves + 1u + wave] = static_cast<uint32_t>(__popcll(m1));
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
        uint32_t s0 = 0, s1 = 0;
        for (uint32_t w = 0; w < kWaves; ++w) {
            const uint32_t a = lds[w], b = lds[kWaves + 1u + w];
            lds[w] = s0;
            lds[kWaves + 1u + w] = s1;
            s0 += a;
            s1 += b;
        }
        lds[kWaves] = s0 ? atomicAdd(c0, s0) : 0u;
        lds[2u * kWaves + 1u] = s1 ? atomicAdd(c1, s1) : 0u;
    }
    __syncthreads();
    r0 = lds[kWaves] + lds[wave] + static_cast<uint32_t>(__popcll(m0 & below));
    r1 = lds[2u * kWaves + 1u] + lds[kWaves + 1u + wave] + static_cast<uint32_t>(__popcll(m1 & below));
}

}  // namespace pt
