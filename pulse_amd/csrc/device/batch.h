// batch.h -- stage kernels of the batched form: many views (each with its
// own scene, camera, resolution and accumulation buffer) traced through one
// sequence of launches.
//
// Layout.  Every view owns a range of slots [pix_base, pix_base + npix) in
// the batch's per-slot arrays (pix_base a multiple of 256); its ray queues,
// hits, path state and shadow queue live in that range, and its queue
// lengths in its own row of `counts`.  A wave never mixes views: the traced
// and shaded queues are cut into chunks of `chunk` rays per view, and a
// chunk index maps to its view through the prefix sum of the per-view chunk
// counts (computed in registers by each wave from the device counters, so a
// captured launch needs no host knowledge of queue lengths).  The view is
// wave-uniform, so its scene pointers sit in scalar registers.
//
// Work distribution.  persistent == 0: one wave per chunk, grid sized for
// the largest possible queue, waves past the end exit.  persistent != 0:
// a fixed grid of waves that take chunks from an atomic counter until the
// queues are drained (every wave reaches the exit test on each iteration).
//
// Regrouping.  The extension queue of trace index k >= 1 is compacted
// (ballot + one atomic per wave) when k % regroup == 0; otherwise the shade
// stage writes each next ray in place, with a dead marker for ended paths,
// and the queue keeps its length.
//
// Per-pixel results do not depend on any of these choices: every pixel's
// path is traced and shaded by the same device functions as the single-view
// kernels, in the same per-pixel order.
#pragma once

#include <type_traits>

#include "pt_device.h"
#include "queue.h"
#include "shade.h"
#include "traverse.h"

namespace pt {

constexpr uint32_t kBatchMaxViews = 256;
constexpr uint32_t kDeadRay = 0xffffffffu;  // pixel field of an in-place dead ray

struct BatchView {
    DevScene sc;
    pulse_camera cam;
    float4 *accum;        // rgba32f running mean, npix entries
    uint32_t pix_base;    // first slot (multiple of 256)
    uint32_t npix;
    uint32_t tiles_x;     // generate tiles per row
    uint32_t tile_base;   // first generate tile of the view
};

struct BatchArgs {
    const BatchView *views;
    uint32_t n_views;
    uint32_t cstride;      // row length of `counts` = 2 * (max_traces + 1)
    float4 *ray_o[2];
    float4 *ray_d[2];
    float4 *hit;
    float *hit_v;
    float4 *state_t;
    float4 *state_c;
    float4 *sh_o;
    float4 *sh_d;
    float4 *sh_c;
    uint32_t *counts;      // [n_views][cstride]: ext[0..max_traces], sh[0..max_traces]
    uint32_t *work;        // persistent work counters: extend k at 2k, shadow k at 2k+1, shade k at 2 max_traces + k
    uint64_t *totals;      // [2]: extension rays, shadow rays (stats)
    const uint32_t *frame; // frame[0]: accumulation count of the first sample
    uint32_t mode, nee_ggx, max_traces, rr_start, horizon_clamp, light_power;
    uint32_t chunk;        // rays per chunk, a multiple of 64
    uint32_t persistent;   // 0: one wave per chunk; else a fixed grid drains the queue
    uint32_t refill;       // persistent single-loop traversal: idle lanes that take new rays
    uint32_t gather;       // persistent: lanes waiting for an instance change before the wave runs it
    uint32_t period;       // persistent: instance changes run at least every `period` wave steps
    int stack_cap;         // traversal stack entries per lane
    uint32_t *spill;       // stack entries below the LDS part, one column per lane of the grid
    uint32_t regroup;      // compact the extension queue of index k when k % regroup == 0
    uint32_t tile_w;       // generate tile width; a tile is tile_w x (64 / tile_w) pixels
};

__device__ __forceinline__ uint32_t uniform_u32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Chunk ranges of the views, held in registers: lane l of group j holds the
// end (exclusive prefix sum + own count) of view 64 j + l's chunk range for
// counter index ci, chunks = ceil(counts[v][ci] / chunk).  Every lane of the
// wave computes it (wave-uniform results).
struct ChunkMap {
    uint32_t end[kBatchMaxViews / 64];
    uint32_t total;
};

__device__ __forceinline__ ChunkMap chunk_map(const BatchArgs &b, uint32_t ci) {
    ChunkMap m;
    const uint32_t lane = __lane_id();
    uint32_t carry = 0;
#pragma unroll
    for (uint32_t j = 0; j < kBatchMaxViews / 64; ++j) {
        const uint32_t v = j * 64u + lane;
        uint32_t x = 0;
        if (j * 64u < b.n_views) {
            x = v < b.n_views ? (b.counts[v * b.cstride + ci] + b.chunk - 1u) / b.chunk : 0u;
            for (int off = 1; off < 64; off <<= 1) {
                uint32_t y = __shfl_up(x, off);
                if (lane >= static_cast<uint32_t>(off)) x += y;
            }
        }
        m.end[j] = carry + x;
        carry += uniform_u32(__shfl(x, 63));
    }
    m.total = carry;
    return m;
}

// View holding chunk g (g < total) and the first chunk of its range: the
// number of views whose range ends at or before g.
__device__ __forceinline__ uint32_t chunk_view(const ChunkMap &m, uint32_t n_views, uint32_t g, uint32_t &start) {
    const uint32_t lane = __lane_id();
    uint32_t v = 0;
#pragma unroll
    for (uint32_t j = 0; j < kBatchMaxViews / 64; ++j) {
        const bool below = j * 64u + lane < n_views && m.end[j] <= g;
        v += static_cast<uint32_t>(__popcll(__ballot(below)));
    }
    start = 0;
    if (v > 0u) {
        const uint32_t u = v - 1u;
        uint32_t e = 0;
#pragma unroll
        for (uint32_t j = 0; j < kBatchMaxViews / 64; ++j)
            if ((u >> 6) == j) e = m.end[j];
        start = uniform_u32(__shfl(e, static_cast<int>(u & 63u)));
    }
    return uniform_u32(v);
}

// Next chunk of a wave: blockIdx-based (one pass) or from the atomic counter.
__device__ __forceinline__ uint32_t next_chunk(const BatchArgs &b, uint32_t wslot, uint32_t wave_global, bool first) {
    if (!b.persistent) return first ? wave_global : 0xffffffffu;
    uint32_t g = 0;
    if (__lane_id() == 0u) g = atomicAdd(&b.work[wslot], 1u);
    return uniform_u32(__shfl(g, 0));
}

// Clears the per-view queue counters and the work counters of one sample.
__global__ void __launch_bounds__(256) k_b_clear(BatchArgs b, uint32_t n_work) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nc = b.n_views * b.cstride;
    if (i < nc) b.counts[i] = 0u;
    else if (i < nc + n_work) b.work[i - nc] = 0u;
}

// Sets frame[0] (a launch argument, so successive frames do not race).
__global__ void k_b_set_frame(uint32_t *frame, uint32_t acc) { frame[0] = acc; }

// generate: primary rays of every view, one tile of 64 pixels per wave.
__global__ void __launch_bounds__(256) k_b_generate(BatchArgs b, uint32_t s) {
    const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t within = threadIdx.x & 63u;
    // view of this tile (tile_base ascending, first view at tile 0)
    uint32_t lo = 0, hi = b.n_views;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (b.views[mid].tile_base <= tile) lo = mid;
        else hi = mid;
    }
    const uint32_t v = uniform_u32(lo);
    const BatchView &bv = b.views[v];
    const uint32_t W = bv.cam.width, H = bv.cam.height;
    const uint32_t tw = b.tile_w, th = 64u / b.tile_w;
    const uint32_t lt = tile - bv.tile_base;
    const uint32_t tx = lt % bv.tiles_x, ty = lt / bv.tiles_x;
    const uint32_t x = tx * tw + within % tw;
    const uint32_t y = ty * th + within / tw;
    const bool valid = x < W && y < H;
    // queue slot in closed form (tiles in order, row-major inside a tile;
    // only the view's last tile row and column are partial): no counter
    const uint32_t tw_v = min(tw, W - min(W, tx * tw)), th_v = min(th, H - min(H, ty * th));
    const uint32_t slot = th * ty * W + th_v * tw * tx + (within / tw) * tw_v + within % tw;
    if (lt == 0u && within == 0u) b.counts[v * b.cstride + 0u] = W * H;
    if (!valid) return;
    const uint32_t pixel = x + y * W;
    V3 o, d;
    primary_ray(bv.cam, x, y, o, d);
    const uint32_t base = bv.pix_base;
    b.ray_o[0][base + slot] = make_float4(o.x, o.y, o.z, __uint_as_float(pixel));
    b.ray_d[0][base + slot] = make_float4(d.x, d.y, d.z, 0.0f);
    const uint32_t rng = pixel_seed(pixel, b.frame[0] + s);
    b.state_t[base + pixel] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(rng));
    b.state_c[base + pixel] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// extend (SHADOW = false: closest hit of extension queue q at trace index k)
// or shadow (SHADOW = true: occlusion of the NEE rays of trace index k; a
// visible light sets the pixel colour that includes its contribution).
template <int STACK, int LOOP, bool SHADOW>
__global__ void __launch_bounds__(64) k_b_trace(BatchArgs b, int q, int k) {
    __shared__ uint32_t smem[STACK * 64];
    const uint32_t ci = SHADOW ? b.max_traces + 1u + static_cast<uint32_t>(k) : static_cast<uint32_t>(k);
    const uint32_t wslot = 2u * static_cast<uint32_t>(k) + (SHADOW ? 1u : 0u);
    const ChunkMap cm = chunk_map(b, ci);
    LdsStack stk{smem + threadIdx.x, 64, STACK};
    const uint32_t lane = __lane_id();
    for (uint32_t g = next_chunk(b, wslot, blockIdx.x, true); g < cm.total; g = next_chunk(b, wslot, blockIdx.x, false)) {
        uint32_t start;
        const uint32_t v = chunk_view(cm, b.n_views, g, start);
        const DevScene sc = b.views[v].sc;
        const uint32_t base = b.views[v].pix_base;
        const uint32_t cnt = b.counts[v * b.cstride + ci];
        const uint32_t first = (g - start) * b.chunk;
        for (uint32_t r = 0; r < b.chunk; r += 64u) {
            const uint32_t i = first + r + lane;
            if (i >= cnt) break;
            const uint32_t slot = base + i;
            if (SHADOW) {
                float4 oa = b.sh_o[slot], da = b.sh_d[slot];
                float t = kTFar, u = 0.0f, vv = 0.0f;
                uint32_t inst = 0, tri = 0;
                bool occluded = trace_any<LOOP, true>(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), t, inst, tri,
                                                      u, vv, da.w, stk);
                if (!occluded) b.state_c[base + __float_as_uint(oa.w)] = b.sh_c[slot];
            } else {
                float4 oa = b.ray_o[q][slot], da = b.ray_d[q][slot];
                float t = kTFar, u = 0.0f, vv = 0.0f;
                uint32_t inst = 0, tri = 0;
                if (__float_as_uint(oa.w) != kDeadRay)
                    trace_any<LOOP, false>(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), t, inst, tri, u, vv, 0.0f,
                                           stk);
                b.hit[slot] = make_float4(t, __uint_as_float(inst), __uint_as_float(tri), u);
                b.hit_v[slot] = vv;
            }
        }
    }
}

// Persistent extend / shadow with per-lane refill: a wave takes a chunk of
// one view's queue (so the scene is wave-uniform), its lanes take the
// chunk's rays one by one as theirs finish (once at least `refill` lanes are
// idle), and the wave takes the next chunk when the current one is drained.
// The top LDS entries of each lane's stack are in LDS, the rest in `spill`.
// Each ray is traversed exactly as by k_b_trace (LOOP 1: the binary BVHs,
// LOOP 2: the 4-wide ones).
//
// `views` is b.views passed again as a read-only, unaliased pointer: the
// scene of the chunk being drained is read through it at every use (scalar
// loads the compiler may repeat) instead of being held in registers across
// the loop.
template <int LDS, bool SHADOW, int LOOP = 1>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LOOP == 2 ? 6 : 1)))
k_b_trace_refill(BatchArgs b, const BatchView *__restrict__ views, int q, int k) {
    __shared__ uint32_t smem[LDS * 64];
    const uint32_t ci = SHADOW ? b.max_traces + 1u + static_cast<uint32_t>(k) : static_cast<uint32_t>(k);
    const uint32_t wslot = 2u * static_cast<uint32_t>(k) + (SHADOW ? 1u : 0u);
    const ChunkMap cm = chunk_map(b, ci);
    LdsStack stk{smem + threadIdx.x, 64, b.stack_cap};
    if (b.stack_cap > LDS) {
        stk.lds_cap = LDS;
        stk.spill = b.spill + blockIdx.x * 64u + threadIdx.x;
        stk.spill_stride = gridDim.x * 64u;
    }
    const uint32_t lane = __lane_id();
    std::conditional_t<LOOP == 2, WideTraversal<SHADOW, 2>, Traversal<SHADOW>> tr;
    bool alive = false;
    uint32_t slot = 0, pixel = 0;
    // wave-uniform: the chunk being drained
    uint32_t base = 0, next = 0, end = 0, vcur = 0;
    uint32_t wave_step = 0;
    for (;;) {
        const DevScene &sc = views[vcur].sc;
        const uint64_t idle = __ballot(!alive);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (next < end && (n_idle >= b.refill || n_idle == 64u)) {
            const uint32_t idx = next + static_cast<uint32_t>(__popcll(idle & ((1ull << lane) - 1ull)));
            if (!alive && idx < end) {
                slot = base + idx;
                if (SHADOW) {
                    float4 oa = b.sh_o[slot], da = b.sh_d[slot];
                    pixel = __float_as_uint(oa.w);
                    alive = tr.begin(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), kTFar, da.w);
                    if (!alive) b.state_c[base + pixel] = b.sh_c[slot];  // empty scene: visible
                } else {
                    float4 oa = b.ray_o[q][slot], da = b.ray_d[q][slot];
                    alive = __float_as_uint(oa.w) != kDeadRay &&
                            tr.begin(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), kTFar, 0.0f);
                    if (!alive) {
                        b.hit[slot] = make_float4(kTFar, 0.0f, 0.0f, 0.0f);
                        b.hit_v[slot] = 0.0f;
                    }
                }
            }
            next = min(next + n_idle, end);
        }
        const uint64_t act = __ballot(alive);
        if (act == 0ull) {
            if (next < end) continue;
            uint32_t g = 0;
            if (lane == 0u) g = atomicAdd(&b.work[wslot], 1u);
            g = uniform_u32(__shfl(g, 0));
            if (g >= cm.total) break;
            uint32_t start;
            const uint32_t v = chunk_view(cm, b.n_views, g, start);
            vcur = v;
            base = views[v].pix_base;
            const uint32_t cnt = b.counts[v * b.cstride + ci];
            next = (g - start) * b.chunk;
            end = min(next + b.chunk, cnt);
            continue;
        }
        // instance changes gathered as in the single-view persistent kernel
        const uint64_t waiting = __ballot(alive && tr.want_advance);
        wave_step += 1u;
        const bool run_advance = waiting == act || static_cast<uint32_t>(__popcll(waiting)) >= b.gather ||
                                 wave_step % b.period == 0u;
        if (alive && !tr.step(sc, stk, run_advance)) {
            alive = false;
            if (SHADOW) {
                if (!tr.occluded) b.state_c[base + pixel] = b.sh_c[slot];
            } else {
                b.hit[slot] = make_float4(tr.rec_t, __uint_as_float(tr.rec_inst), __uint_as_float(tr.rec_tri), tr.rec_u);
                b.hit_v[slot] = tr.rec_v;
            }
        }
    }
}

// One path vertex of the batch: shades extension ray `i` (in slot base + i)
// if valid; the path state is updated in place.
__device__ __forceinline__ ShadeOut b_shade_one(const BatchArgs &b, const DevScene &sc, const FrameArgs &fa, int q,
                                                int k, uint32_t base, uint32_t i, bool valid, uint32_t &pixel) {
    ShadeOut so;
    pixel = kDeadRay;
    if (valid) {
        const uint32_t slot = base + i;
        float4 oa = b.ray_o[q][slot], da = b.ray_d[q][slot];
        pixel = __float_as_uint(oa.w);
        float4 ha = b.hit[slot];
        if (pixel != kDeadRay && ha.x < kTFar) {
            float4 st = b.state_t[base + pixel], sc4 = b.state_c[base + pixel];
            V3 T = mk(st.x, st.y, st.z), C = mk(sc4.x, sc4.y, sc4.z);
            uint32_t rng = __float_as_uint(st.w);
            so = shade_ray(sc, fa, k, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), ha, b.hit_v[slot], T, C, rng);
            b.state_t[base + pixel] = make_float4(T.x, T.y, T.z, __uint_as_float(rng));
            b.state_c[base + pixel] = make_float4(C.x, C.y, C.z, 0.0f);
        }
    }
    return so;
}

// shade: one bounce for extension queue q at trace index k, 4 waves per
// workgroup.  Persistent: the workgroup takes one chunk at a time (all its
// waves on one view) and appends the chunk's next and shadow rays with one
// atomic per 256 rays; otherwise each wave on its own chunk.
__global__ void __launch_bounds__(256) k_b_shade(BatchArgs b, int q, int k) {
    __shared__ uint32_t s_chunk;
    __shared__ uint32_t append_lds[2 * (256 / 64 + 1)];
    const uint32_t ci = static_cast<uint32_t>(k);
    const ChunkMap cm = chunk_map(b, ci);
    const uint32_t lane = __lane_id();
    const uint32_t wave_global = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t wslot = 2u * b.max_traces + static_cast<uint32_t>(k);
    const bool compact = ((static_cast<uint32_t>(k) + 1u) % b.regroup) == 0u;
    FrameArgs fa;
    fa.mode = b.mode;
    fa.nee_ggx = b.nee_ggx;
    fa.max_traces = b.max_traces;
    fa.rr_start = b.rr_start;
    fa.horizon_clamp = b.horizon_clamp;
    fa.light_power = b.light_power;
    fa.gi = 0;
    fa.accumulation_count = 0;
    if (b.persistent) {
        for (;;) {
            if (threadIdx.x == 0u) s_chunk = atomicAdd(&b.work[wslot], 1u);
            __syncthreads();
            const uint32_t g = s_chunk;
            __syncthreads();  // every thread has it before the next claim
            if (g >= cm.total) break;
            uint32_t start;
            const uint32_t v = chunk_view(cm, b.n_views, g, start);
            const DevScene sc = b.views[v].sc;
            const uint32_t base = b.views[v].pix_base;
            uint32_t *row = b.counts + v * b.cstride;
            const uint32_t cnt = row[ci];
            const uint32_t first = (g - start) * b.chunk;
            if (!compact && first == 0u && threadIdx.x == 0u) row[ci + 1u] = cnt;  // in place: same length
            for (uint32_t r = 0; r < b.chunk && first + r < cnt; r += 256u) {
                const uint32_t i = first + r + threadIdx.x;
                const bool valid = r + threadIdx.x < b.chunk && i < cnt;
                uint32_t pixel;
                const ShadeOut so = b_shade_one(b, sc, fa, q, k, base, i, valid, pixel);
                uint32_t rs, ss;
                block_append2<256>(compact && so.emit_ray, &row[ci + 1u], so.emit_shadow, &row[b.max_traces + 1u + ci],
                                   append_lds, rs, ss);
                if (compact) {
                    if (so.emit_ray) {
                        b.ray_o[q ^ 1][base + rs] =
                            make_float4(so.next_o.x, so.next_o.y, so.next_o.z, __uint_as_float(pixel));
                        b.ray_d[q ^ 1][base + rs] = make_float4(so.next_d.x, so.next_d.y, so.next_d.z, 0.0f);
                    }
                } else if (valid) {
                    const uint32_t tag = so.emit_ray ? pixel : kDeadRay;
                    b.ray_o[q ^ 1][base + i] = make_float4(so.next_o.x, so.next_o.y, so.next_o.z, __uint_as_float(tag));
                    b.ray_d[q ^ 1][base + i] = make_float4(so.next_d.x, so.next_d.y, so.next_d.z, 0.0f);
                }
                if (so.emit_shadow) {
                    b.sh_o[base + ss] = make_float4(so.sh_o.x, so.sh_o.y, so.sh_o.z, __uint_as_float(pixel));
                    b.sh_d[base + ss] = make_float4(so.sh_d.x, so.sh_d.y, so.sh_d.z, so.sh_thr);
                    b.sh_c[base + ss] = make_float4(so.sh_c.x, so.sh_c.y, so.sh_c.z, 0.0f);
                }
            }
        }
        return;
    }
    for (uint32_t g = next_chunk(b, wslot, wave_global, true); g < cm.total;
         g = next_chunk(b, wslot, wave_global, false)) {
        uint32_t start;
        const uint32_t v = chunk_view(cm, b.n_views, g, start);
        const DevScene sc = b.views[v].sc;
        const uint32_t base = b.views[v].pix_base;
        uint32_t *row = b.counts + v * b.cstride;
        const uint32_t cnt = row[ci];
        const uint32_t first = (g - start) * b.chunk;
        if (!compact && first == 0u && lane == 0u) row[ci + 1u] = cnt;  // in place: same length
        for (uint32_t r = 0; r < b.chunk && first + r < cnt; r += 64u) {
            const uint32_t i = first + r + lane;
            const bool valid = i < cnt;
            uint32_t pixel;
            const ShadeOut so = b_shade_one(b, sc, fa, q, k, base, i, valid, pixel);
            if (compact) {
                const uint32_t rs = wave_append(so.emit_ray, &row[ci + 1u]);
                if (so.emit_ray) {
                    b.ray_o[q ^ 1][base + rs] = make_float4(so.next_o.x, so.next_o.y, so.next_o.z, __uint_as_float(pixel));
                    b.ray_d[q ^ 1][base + rs] = make_float4(so.next_d.x, so.next_d.y, so.next_d.z, 0.0f);
                }
            } else if (valid) {
                const uint32_t tag = so.emit_ray ? pixel : kDeadRay;
                b.ray_o[q ^ 1][base + i] = make_float4(so.next_o.x, so.next_o.y, so.next_o.z, __uint_as_float(tag));
                b.ray_d[q ^ 1][base + i] = make_float4(so.next_d.x, so.next_d.y, so.next_d.z, 0.0f);
            }
            const uint32_t ss = wave_append(so.emit_shadow, &row[b.max_traces + 1u + ci]);
            if (so.emit_shadow) {
                b.sh_o[base + ss] = make_float4(so.sh_o.x, so.sh_o.y, so.sh_o.z, __uint_as_float(pixel));
                b.sh_d[base + ss] = make_float4(so.sh_d.x, so.sh_d.y, so.sh_d.z, so.sh_thr);
                b.sh_c[base + ss] = make_float4(so.sh_c.x, so.sh_c.y, so.sh_c.z, 0.0f);
            }
        }
    }
}

// accumulate: running mean of every view's pixels (path_tracer.wgsl:116-123),
// and the sample's ray counts into `totals`.
__global__ void __launch_bounds__(256) k_b_accumulate(BatchArgs b, uint32_t s) {
    const uint32_t slot = blockIdx.x * 256u + threadIdx.x;
    uint32_t lo = 0, hi = b.n_views;  // view with pix_base <= slot (block-uniform: bases are multiples of 256)
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (b.views[mid].pix_base <= slot) lo = mid;
        else hi = mid;
    }
    const uint32_t v = uniform_u32(lo);
    const BatchView &bv = b.views[v];
    if (blockIdx.x == 0u && threadIdx.x < 64u) {
        // ray counts of this sample: lane j sums trace index j over views
        uint64_t ext = 0, sh = 0;
        if (threadIdx.x < b.max_traces) {
            for (uint32_t u = 0; u < b.n_views; ++u) {
                ext += b.counts[u * b.cstride + threadIdx.x];
                sh += b.counts[u * b.cstride + b.max_traces + 1u + threadIdx.x];
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            ext += __shfl_down(ext, off);
            sh += __shfl_down(sh, off);
        }
        if (threadIdx.x == 0u) {
            atomicAdd(reinterpret_cast<unsigned long long *>(&b.totals[0]), static_cast<unsigned long long>(ext));
            atomicAdd(reinterpret_cast<unsigned long long *>(&b.totals[1]), static_cast<unsigned long long>(sh));
        }
    }
    const uint32_t p = slot - bv.pix_base;
    if (p >= bv.npix) return;
    const uint32_t n = b.frame[0] + s;
    float4 c = b.state_c[slot];
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if (n > 0u) {
        float4 old = bv.accum[p];
        ox = old.x;
        oy = old.y;
        oz = old.z;
    }
    float weight = 1.0f / (static_cast<float>(n) + 1.0f);
    float omw = 1.0f - weight;
    bv.accum[p] = make_float4(ox * omw + c.x * weight, oy * omw + c.y * weight, oz * omw + c.z * weight, 1.0f);
}

}  // namespace pt
