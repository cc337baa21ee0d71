// wavefront.h -- the stage kernels of the wavefront (ray-queue) path tracer.
//
// One path per pixel per sample.  The reference runs the whole path in one
// invocation per pixel (path_trace, path_tracer.wgsl:58-124); here each
// bounce is split into stages that run over compacted queues:
//   generate  -> primary rays (8x8 pixel tiles per wave), path state init
//   extend    -> closest hit for every queued ray            (traversal)
//   shade     -> emission, NEE setup, BSDF sampling, RR;     emits shadow
//                rays and next-bounce rays into compacted queues
//   shadow    -> occlusion test of the NEE rays              (traversal)
//   accumulate-> running mean into the rgba32f target
// Queues are compacted per wave with a ballot and one atomic per wave.  The
// per-pixel order of colour updates is the reference's (emission/NEE of
// bounce k is folded in before bounce k+1 is shaded), so each pixel's result
// does not depend on queue order.
#pragma once

#include <type_traits>

#include "pt_device.h"
#include "queue.h"
#include "shade.h"
#include "traverse.h"

namespace pt {

constexpr int kShadeBlock = 256;

// Wavefront buffers (device pointers), shared by all stages.
struct WfBuffers {
    // extension ray queues (ping-pong): {o.xyz, pixel bits}, {d.xyz, 0}
    float4 *ray_o[2];
    float4 *ray_d[2];
    // closest hit per queue slot: {t, inst bits, tri bits, u}, v
    float4 *hit;
    float *hit_v;
    // per-pixel path state: {throughput.xyz, rng bits}, {color.xyz, 0}
    float4 *state_t;
    float4 *state_c;
    // shadow queue: {o.xyz, pixel bits}, {d.xyz, threshold}, {color if visible}
    float4 *sh_o;
    float4 *sh_d;
    float4 *sh_c;
    // counters: ext[k] = rays traced at trace index k; sh[k] = shadow rays
    uint32_t *ext_count;  // [kMaxTraces + 1]
    uint32_t *sh_count;   // [kMaxTraces + 1]
    // rays taken by the persistent extend / shadow launches from part p of
    // the queue at trace index k: [(k * kQueueParts + p) * kHeadStride]
    uint32_t *ext_head;
    uint32_t *sh_head;
    // optional record of the paths for the adjoint (null: none): at
    // r = (rec_base + pixel) * rec_traces + trace index, the closest hit
    // {t, inst bits, tri bits, u}, its v, and 1 / 0 for an unoccluded /
    // occluded shadow ray (rec_base = sample * pixels)
    float4 *rec_hit = nullptr;
    float *rec_v = nullptr;
    uint8_t *rec_vis = nullptr;
    uint32_t rec_base = 0, rec_traces = 0;
};

// generate: primary rays (path_tracer.wgsl:60-75) in 8x8 pixel tiles.
__global__ void __launch_bounds__(kShadeBlock) k_generate(WfBuffers wb, FrameArgs fa) {
    const uint32_t W = fa.cam.width, H = fa.cam.height;
    const uint32_t tiles_x = (W + 7u) / 8u;
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t tile = s >> 6, within = s & 63u;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    uint32_t x = tx * 8u + (within & 7u);
    uint32_t y = ty * 8u + (within >> 3);
    bool valid = x < W && y < H;
    // queue slot: tile after tile, row-major inside a tile (only the last
    // tile row and column are partial: the rows above hold 8 W pixels, the
    // tiles to the left th * 8 each), so no counter is contended
    const uint32_t tw = min(8u, W - tx * 8u), th = min(8u, H - min(H, ty * 8u));
    const uint32_t slot = 8u * ty * W + th * 8u * tx + (within >> 3) * tw + (within & 7u);
    if (s == 0u) wb.ext_count[0] = W * H;
    if (!valid) return;
    uint32_t pixel = x + y * W;
    V3 o, d;
    primary_ray(fa.cam, x, y, o, d);
    wb.ray_o[0][slot] = make_float4(o.x, o.y, o.z, __uint_as_float(pixel));
    wb.ray_d[0][slot] = make_float4(d.x, d.y, d.z, 0.0f);
    uint32_t rng = pixel_seed(pixel, fa.accumulation_count);
    wb.state_t[pixel] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(rng));
    wb.state_c[pixel] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// extend: closest hit for queue `q` at trace index k.
template <int BLOCK, int STACK, int LOOP>
__global__ void __launch_bounds__(BLOCK) k_extend(DevScene sc, WfBuffers wb, int q, int k) {
    __shared__ uint32_t smem[STACK * BLOCK];
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= wb.ext_count[k]) return;
    LdsStack stk{smem + threadIdx.x, BLOCK, STACK};
    float4 oa = wb.ray_o[q][i], da = wb.ray_d[q][i];
    V3 o = mk(oa.x, oa.y, oa.z), d = mk(da.x, da.y, da.z);
    float t = kTFar, u = 0.0f, v = 0.0f;
    uint32_t inst = 0, tri = 0;
    trace_any<LOOP, false>(sc, o, d, t, inst, tri, u, v, 0.0f, stk);
    wb.hit[i] = make_float4(t, __uint_as_float(inst), __uint_as_float(tri), u);
    wb.hit_v[i] = v;
}

// shade: one bounce of the path loop for every ray of queue `q` (trace index
// k).  Next-bounce rays go to queue q^1 / counter k+1, shadow rays to the
// shadow queue / counter k.
__global__ void __launch_bounds__(kShadeBlock) k_shade(DevScene sc, WfBuffers wb, FrameArgs fa, int q, int k) {
    __shared__ uint32_t append_lds[2 * (kShadeBlock / 64 + 1)];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < wb.ext_count[k];
    ShadeOut so;
    uint32_t pixel = 0;
    if (valid) {
        float4 oa = wb.ray_o[q][i], da = wb.ray_d[q][i];
        float4 ha = wb.hit[i];
        pixel = __float_as_uint(oa.w);
        if (wb.rec_hit) {
            const uint32_t r = (wb.rec_base + pixel) * wb.rec_traces + static_cast<uint32_t>(k);
            wb.rec_hit[r] = ha;
            wb.rec_v[r] = wb.hit_v[i];
        }
        if (ha.x < kTFar) {  // a miss ends the path (path_tracer.wgsl:83-86)
            float4 st = wb.state_t[pixel], sc4 = wb.state_c[pixel];
            V3 T = mk(st.x, st.y, st.z), C = mk(sc4.x, sc4.y, sc4.z);
            uint32_t rng = __float_as_uint(st.w);
            so = shade_ray(sc, fa, k, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), ha, wb.hit_v[i], T, C, rng);
            wb.state_t[pixel] = make_float4(T.x, T.y, T.z, __uint_as_float(rng));
            wb.state_c[pixel] = make_float4(C.x, C.y, C.z, 0.0f);
        }
    }
    uint32_t rs, ss;
    block_append2<kShadeBlock>(so.emit_ray, &wb.ext_count[k + 1], so.emit_shadow, &wb.sh_count[k], append_lds, rs, ss);
    if (so.emit_ray) {
        wb.ray_o[q ^ 1][rs] = make_float4(so.next_o.x, so.next_o.y, so.next_o.z, __uint_as_float(pixel));
        wb.ray_d[q ^ 1][rs] = make_float4(so.next_d.x, so.next_d.y, so.next_d.z, 0.0f);
    }
    if (so.emit_shadow) {
        wb.sh_o[ss] = make_float4(so.sh_o.x, so.sh_o.y, so.sh_o.z, __uint_as_float(pixel));
        wb.sh_d[ss] = make_float4(so.sh_d.x, so.sh_d.y, so.sh_d.z, so.sh_thr);
        wb.sh_c[ss] = make_float4(so.sh_c.x, so.sh_c.y, so.sh_c.z, 0.0f);
    }
}

// shadow: occlusion test of the NEE rays of trace index k; a visible light
// replaces the pixel colour by the one that includes its contribution.
template <int BLOCK, int STACK, int LOOP>
__global__ void __launch_bounds__(BLOCK) k_shadow(DevScene sc, WfBuffers wb, int k) {
    __shared__ uint32_t smem[STACK * BLOCK];
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= wb.sh_count[k]) return;
    LdsStack stk{smem + threadIdx.x, BLOCK, STACK};
    float4 oa = wb.sh_o[i], da = wb.sh_d[i];
    V3 o = mk(oa.x, oa.y, oa.z), d = mk(da.x, da.y, da.z);
    float t = kTFar, u = 0.0f, v = 0.0f;
    uint32_t inst = 0, tri = 0;
    bool occluded = trace_any<LOOP, true>(sc, o, d, t, inst, tri, u, v, da.w, stk);
    if (!occluded) {
        float4 c = wb.sh_c[i];
        wb.state_c[__float_as_uint(oa.w)] = c;
    }
    if (wb.rec_vis) wb.rec_vis[(wb.rec_base + __float_as_uint(oa.w)) * wb.rec_traces + k] = occluded ? 0u : 1u;
}

// Persistent extend (SHADOW = false, queue q at trace index k) or shadow
// (SHADOW = true) over a fixed grid of one-wave workgroups.  A lane whose ray
// is done takes the next one from the queue (one atomic per wave for all
// idle lanes) once at least `refill` lanes of its wave are idle, so a wave
// does not idle behind its longest ray.  Each ray is traversed exactly as by
// k_extend / k_shadow (same steps, same results); only the assignment of
// rays to lanes changes.
//
// Instance changes (the longest code path of a step) are gathered: a lane
// whose next operation is one waits until at least `gather` lanes of the
// wave are waiting for one, no lane of the wave has anything else to do, or
// the wave's step count is a multiple of `period`.
//
// The top LDS_ENTRIES entries of each lane's stack live in LDS, the rest (up
// to `capacity`) in `spill`, one column of capacity - LDS_ENTRIES words per
// lane of the grid.  LOOP 1 walks the reference's binary BVHs, LOOP 2 the
// 4-wide ones (wide.h).
template <int LDS_ENTRIES, bool SHADOW, int TPS, int LOOP = 1>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SHADOW ? 7 : 6)))
k_trace_queue(DevScene sc, WfBuffers wb, int q, int k, uint32_t refill, uint32_t gather, uint32_t period,
              int capacity, uint32_t *spill) {
    __shared__ uint32_t smem[LDS_ENTRIES * 64];
    LdsStack stk{smem + threadIdx.x, 64, capacity};
    if (capacity > LDS_ENTRIES) {
        stk.lds_cap = LDS_ENTRIES;
        stk.spill = spill + blockIdx.x * 64u + threadIdx.x;
        stk.spill_stride = gridDim.x * 64u;
    }
    const uint32_t count = SHADOW ? wb.sh_count[k] : wb.ext_count[k];
    uint32_t *heads = (SHADOW ? wb.sh_head : wb.ext_head) + k * (kQueueParts * kHeadStride);
    const uint32_t lane = __lane_id();
    // wave-uniform: the queue part rays are taken from (the block's XCD
    // first), and how many parts were found drained
    uint32_t part = blockIdx.x % kQueueParts, drained = 0;
    std::conditional_t<LOOP == 2, WideTraversal<SHADOW, TPS>, Traversal<SHADOW, TPS>> tr;
    bool alive = false;
    bool exhausted = false;  // wave-uniform: the queue has no rays left
    uint32_t wave_step = 0;
    uint32_t slot = 0, pixel = 0;
    for (;;) {
        const uint64_t idle = __ballot(!alive);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (!exhausted && (n_idle >= refill || n_idle == 64u)) {
            // part p holds slots [count * p / parts, count * (p + 1) / parts)
            const uint32_t lo = static_cast<uint32_t>(static_cast<uint64_t>(count) * part / kQueueParts);
            const uint32_t hi = static_cast<uint32_t>(static_cast<uint64_t>(count) * (part + 1u) / kQueueParts);
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(&heads[part * kHeadStride], n_idle);
            base = lo + __shfl(base, 0);
            if (base + n_idle >= hi) {
                // this part is drained: later refills take from the next one
                part = (part + 1u) % kQueueParts;
                if (++drained == kQueueParts) exhausted = true;
            }
            if (!alive) {
                const uint32_t idx = base + static_cast<uint32_t>(__popcll(idle & ((1ull << lane) - 1ull)));
                if (idx < hi) {
                    slot = idx;
                    if (SHADOW) {
                        float4 oa = wb.sh_o[idx], da = wb.sh_d[idx];
                        pixel = __float_as_uint(oa.w);
                        alive = tr.begin(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), kTFar, da.w);
                        if (!alive) {
                            wb.state_c[pixel] = wb.sh_c[idx];  // empty scene: visible
                            if (wb.rec_vis) wb.rec_vis[(wb.rec_base + pixel) * wb.rec_traces + k] = 1u;
                        }
                    } else {
                        float4 oa = wb.ray_o[q][idx], da = wb.ray_d[q][idx];
                        alive = tr.begin(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), kTFar, 0.0f);
                        if (!alive) {
                            wb.hit[idx] = make_float4(kTFar, 0.0f, 0.0f, 0.0f);
                            wb.hit_v[idx] = 0.0f;
                        }
                    }
                }
            }
        }
        const uint64_t act = __ballot(alive);
        if (act == 0ull) {
            if (exhausted) break;
            continue;
        }
        const uint64_t waiting = __ballot(alive && tr.want_advance);
        wave_step += 1u;
        const bool run_advance = waiting == act || static_cast<uint32_t>(__popcll(waiting)) >= gather ||
                                 wave_step % period == 0u;
        if (alive && !tr.step(sc, stk, run_advance)) {
            alive = false;
            if (SHADOW) {
                if (!tr.occluded) wb.state_c[pixel] = wb.sh_c[slot];
                if (wb.rec_vis) wb.rec_vis[(wb.rec_base + pixel) * wb.rec_traces + k] = tr.occluded ? 0u : 1u;
            } else {
                wb.hit[slot] = make_float4(tr.rec_t, __uint_as_float(tr.rec_inst), __uint_as_float(tr.rec_tri), tr.rec_u);
                wb.hit_v[slot] = tr.rec_v;
            }
        }
    }
}

// Work census of one queue for the 4-wide traversal: walks every ray of
// queue q at trace index k (the shadow queue if SHADOW) as the trace kernels
// do, writes nothing back, and adds {rays, wide node visits, triangle tests,
// instances entered} to out[0..3] (one atomic per counter per wave).
template <int STACK, bool SHADOW>
__global__ void __launch_bounds__(64) k_census(DevScene sc, WfBuffers wb, int q, int k, unsigned long long *out) {
    __shared__ uint32_t smem[STACK * 64];
    LdsStack stk{smem + threadIdx.x, 64, STACK};
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    const uint32_t count = SHADOW ? wb.sh_count[k] : wb.ext_count[k];
    uint32_t c[4] = {0u, 0u, 0u, 0u};
    if (i < count) {
        WideTraversal<SHADOW, 2, true> tr;
        const float4 oa = SHADOW ? wb.sh_o[i] : wb.ray_o[q][i];
        const float4 da = SHADOW ? wb.sh_d[i] : wb.ray_d[q][i];
        if (tr.begin(sc, mk(oa.x, oa.y, oa.z), mk(da.x, da.y, da.z), kTFar, SHADOW ? da.w : 0.0f)) {
            while (tr.step(sc, stk)) {
            }
            c[1] = static_cast<uint32_t>(tr.iter);
            c[2] = tr.n_tri;
            c[3] = tr.n_inst;
        }
        c[0] = 1u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t v = c[j];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (threadIdx.x == 0 && v) atomicAdd(&out[j], static_cast<unsigned long long>(v));
    }
}

// G-buffer GI pass (gi.wgsl:56-126), the part before the first trace: per
// pixel, from its G-buffer sample, a cosine scatter ray and the NEE shadow
// ray.  Sample 0 seeds the RNG with pixel_index * 5817321; later samples go
// on from where the previous sample's path left it.
struct GBufferArgs {
    const float4 *position, *normal, *base_color, *emissive;
    uint32_t npix;
};
__global__ void __launch_bounds__(kShadeBlock) k_gi_generate(DevScene sc, WfBuffers wb, FrameArgs fa, GBufferArgs g,
                                                             uint32_t s) {
    __shared__ uint32_t append_lds[2 * (kShadeBlock / 64 + 1)];
    const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pixel < g.npix;
    ShadeOut so;
    if (valid) {
        uint32_t rng = s == 0u ? pixel * 5817321u : __float_as_uint(wb.state_t[pixel].w);
        const float4 p4 = g.position[pixel], n4 = g.normal[pixel], b4 = g.base_color[pixel], e4 = g.emissive[pixel];
        const V3 p = mk(p4.x, p4.y, p4.z), n = mk(n4.x, n4.y, n4.z);
        Material m;
        m.base = mk(b4.x, b4.y, b4.z);
        m.emissive = mk(e4.x, e4.y, e4.z);
        m.roughness = 0.5f;
        m.reflectance = 0.5f;
        m.metallic = 0.0f;
        float a0 = rand_f(rng);
        float a1 = rand_f(rng);
        V3 scatter = sample_cosine_hemisphere(n, a0, a1);
        V3 dl_vis = mk(0.0f, 0.0f, 0.0f);
        const bool lights = sc.light_count > 0u;
        if (lights) {
            NeeSample ns = nee_sample(sc, p, n, m, neg(n), false, rng, fa.light_power != 0u);
            so.emit_shadow = true;
            so.sh_o = ns.origin;
            so.sh_d = ns.dir;
            so.sh_thr = ns.threshold;
            dl_vis = ns.contrib;
        }
        so.sh_c = add(m.emissive, dl_vis);
        V3 C = lights ? add(m.emissive, mk(0.0f, 0.0f, 0.0f)) : so.sh_c;
        V3 T = m.base;
        wb.state_t[pixel] = make_float4(T.x, T.y, T.z, __uint_as_float(rng));
        wb.state_c[pixel] = make_float4(C.x, C.y, C.z, 0.0f);
        so.emit_ray = fa.max_traces > 0u;
        so.next_o = add(p, scale(0.001f, n));
        so.next_d = scatter;
    }
    uint32_t rs, ss;
    block_append2<kShadeBlock>(so.emit_ray, &wb.ext_count[0], so.emit_shadow, &wb.sh_count[kMaxTraces], append_lds, rs,
                               ss);
    if (so.emit_ray) {
        wb.ray_o[0][rs] = make_float4(so.next_o.x, so.next_o.y, so.next_o.z, __uint_as_float(pixel));
        wb.ray_d[0][rs] = make_float4(so.next_d.x, so.next_d.y, so.next_d.z, 0.0f);
    }
    if (so.emit_shadow) {
        wb.sh_o[ss] = make_float4(so.sh_o.x, so.sh_o.y, so.sh_o.z, __uint_as_float(pixel));
        wb.sh_d[ss] = make_float4(so.sh_d.x, so.sh_d.y, so.sh_d.z, so.sh_thr);
        wb.sh_c[ss] = make_float4(so.sh_c.x, so.sh_c.y, so.sh_c.z, 0.0f);
    }
}

// G-buffer GI pass output: the running sum of colour * (1 / spp), alpha 1
// (gi.wgsl:122, 130).
__global__ void __launch_bounds__(kShadeBlock) k_gi_accumulate(WfBuffers wb, float4 *out, uint32_t npix, uint32_t s,
                                                               float spp_inv) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float4 c = wb.state_c[p];
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if (s > 0u) {
        const float4 o = out[p];
        ox = o.x;
        oy = o.y;
        oz = o.z;
    }
    out[p] = make_float4(ox + c.x * spp_inv, oy + c.y * spp_inv, oz + c.z * spp_inv, 1.0f);
}

// accumulate: running mean (path_tracer.wgsl:116-123).
__global__ void __launch_bounds__(kShadeBlock) k_accumulate(WfBuffers wb, float4 *accum, uint32_t npix, uint32_t n) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float4 c = wb.state_c[p];
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if (n > 0u) {
        float4 old = accum[p];
        ox = old.x;
        oy = old.y;
        oz = old.z;
    }
    float weight = 1.0f / (static_cast<float>(n) + 1.0f);
    float omw = 1.0f - weight;
    accum[p] = make_float4(ox * omw + c.x * weight, oy * omw + c.y * weight, oz * omw + c.z * weight, 1.0f);
}

// Debug/parity entry: closest hit per ray.
template <int BLOCK, int STACK, int LOOP>
__global__ void __launch_bounds__(BLOCK) k_trace_rays(DevScene sc, const pulse_ray *rays, pulse_hit *hits, uint32_t n) {
    __shared__ uint32_t smem[STACK * BLOCK];
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    LdsStack stk{smem + threadIdx.x, BLOCK, STACK};
    pulse_ray r = rays[i];
    V3 o = mk(r.origin[0], r.origin[1], r.origin[2]), d = mk(r.dir[0], r.dir[1], r.dir[2]);
    float t = r.t_max, u = 0.0f, v = 0.0f;
    uint32_t inst = 0, tri = 0;
    trace_any<LOOP, false>(sc, o, d, t, inst, tri, u, v, 0.0f, stk);
    pulse_hit h;
    h.t = t;
    h.instance_index = inst;
    h.triangle_index = tri;
    h.u = u;
    h.v = v;
    hits[i] = h;
}

template <int BLOCK, int STACK, int LOOP>
__global__ void __launch_bounds__(BLOCK) k_shadow_rays(DevScene sc, const pulse_ray *rays, uint32_t *occ, uint32_t n) {
    __shared__ uint32_t smem[STACK * BLOCK];
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    LdsStack stk{smem + threadIdx.x, BLOCK, STACK};
    pulse_ray r = rays[i];
    V3 o = mk(r.origin[0], r.origin[1], r.origin[2]), d = mk(r.dir[0], r.dir[1], r.dir[2]);
    float t = kTFar, u = 0.0f, v = 0.0f;
    uint32_t inst = 0, tri = 0;
    occ[i] = trace_any<LOOP, true>(sc, o, d, t, inst, tri, u, v, r.t_max, stk) ? 1u : 0u;
}

}  // namespace pt
