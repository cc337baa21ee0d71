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

#include "pt_device.h"
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
};
constexpr int kMaxTraces = 64;

// Traversal variants: 0 = the reference's nested TLAS/BLAS loops,
// 1 = single loop over both levels (traverse.h).  Same results.
template <int LOOP, bool ANY_HIT>
__device__ __forceinline__ bool trace_any(const DevScene &sc, V3 o, V3 d, float &t, uint32_t &inst, uint32_t &tri,
                                          float &u, float &v, float threshold, LdsStack &stk) {
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

// generate: primary rays (path_tracer.wgsl:60-75) in 8x8 pixel tiles.
__global__ void __launch_bounds__(kShadeBlock) k_generate(WfBuffers wb, FrameArgs fa) {
    const uint32_t W = fa.cam.width, H = fa.cam.height;
    const uint32_t tiles_x = (W + 7u) / 8u;
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t tile = s >> 6, within = s & 63u;
    uint32_t x = (tile % tiles_x) * 8u + (within & 7u);
    uint32_t y = (tile / tiles_x) * 8u + (within >> 3);
    bool valid = x < W && y < H;
    uint32_t slot = wave_append(valid, &wb.ext_count[0]);
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
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < wb.ext_count[k];
    bool emit_ray = false, emit_shadow = false;
    uint32_t pixel = 0;
    V3 next_o{}, next_d{}, sh_o{}, sh_d{}, sh_c{};
    float sh_thr = 0.0f;
    V3 T{}, C{};
    uint32_t rng = 0;
    if (valid) {
        float4 oa = wb.ray_o[q][i], da = wb.ray_d[q][i];
        float4 ha = wb.hit[i];
        pixel = __float_as_uint(oa.w);
        V3 o = mk(oa.x, oa.y, oa.z), d = mk(da.x, da.y, da.z);
        float t = ha.x;
        if (t < kTFar) {  // a miss ends the path (path_tracer.wgsl:83-86)
            float4 st = wb.state_t[pixel], sc4 = wb.state_c[pixel];
            T = mk(st.x, st.y, st.z);
            rng = __float_as_uint(st.w);
            C = mk(sc4.x, sc4.y, sc4.z);
            HitInfo h = shade_hit(sc, o, d, t, __float_as_uint(ha.y), __float_as_uint(ha.z), ha.w, wb.hit_v[i]);
            Material m = load_material(sc.materials[h.material_index]);
            V3 wo = neg(d);
            const bool last = static_cast<uint32_t>(k) + 1u >= fa.max_traces;
            const bool lights = sc.light_count > 0u;
            if (fa.mode == PULSE_MODE_REF_NEE_LAMBERT) {
                if (k == 0) {
                    // first hit stands in for the G-buffer sample
                    // (path_tracer.wgsl:143-153)
                    float a0 = rand_f(rng);
                    float a1 = rand_f(rng);
                    V3 scatter = sample_cosine_hemisphere(h.normal, a0, a1);
                    V3 dl_vis = mk(0.0f, 0.0f, 0.0f);
                    if (lights) {
                        NeeSample ns = nee_sample(sc, h.position, h.normal, m, wo, false, rng);
                        emit_shadow = true;
                        sh_o = ns.origin;
                        sh_d = ns.dir;
                        sh_thr = ns.threshold;
                        dl_vis = ns.contrib;
                    }
                    sh_c = add(m.emissive, dl_vis);
                    C = add(m.emissive, mk(0.0f, 0.0f, 0.0f));
                    if (!lights) C = sh_c;
                    T = m.base;
                    if (!last) {
                        emit_ray = true;
                        next_o = add(h.position, scale(0.001f, h.normal));
                        next_d = scatter;
                    }
                } else {
                    // loop body, path_tracer.wgsl:175-190
                    if (lights) {
                        NeeSample ns = nee_sample(sc, h.position, h.normal, m, wo, false, rng);
                        emit_shadow = true;
                        sh_o = ns.origin;
                        sh_d = ns.dir;
                        sh_thr = ns.threshold;
                        sh_c = add(C, mul(T, ns.contrib));
                    }
                    // without lights the reference adds throughput * 0 too
                    C = add(C, mul(T, mk(0.0f, 0.0f, 0.0f)));
                    if (!last) {
                        T = mul(T, m.base);
                        bool alive = true;
                        if (static_cast<uint32_t>(k - 1) >= fa.rr_start) {
                            float p = max3(T);
                            if (rand_f(rng) > p) alive = false;
                            else T = scale(1.0f / p, T);
                        }
                        if (alive) {
                            float b0 = rand_f(rng);
                            float b1 = rand_f(rng);
                            emit_ray = true;
                            next_d = sample_cosine_hemisphere(h.normal, b0, b1);
                            next_o = add(h.position, scale(0.001f, h.normal));
                        }
                    }
                }
            } else {
                if (fa.mode == PULSE_MODE_REF_GGX || k == 0) C = add(C, mul(T, m.emissive));
                if (fa.mode == PULSE_MODE_GGX_NEE && lights) {
                    NeeSample ns = nee_sample(sc, h.position, h.normal, m, wo, fa.nee_ggx != 0u, rng);
                    emit_shadow = true;
                    sh_o = ns.origin;
                    sh_d = ns.dir;
                    sh_thr = ns.threshold;
                    sh_c = add(C, mul(T, ns.contrib));
                    C = add(C, mul(T, mk(0.0f, 0.0f, 0.0f)));
                }
                if (!last) {
                    Sample smp = importance_sample_ggx_d(h.normal, wo, m, rng);
                    bool alive = !(fa.horizon_clamp && below_horizon(smp));
                    if (alive) T = mul(T, smp.reflectance);
                    if (alive && static_cast<uint32_t>(k) >= fa.rr_start) {
                        float p = max3(T);
                        if (rand_f(rng) > p) alive = false;
                        else T = scale(1.0f / p, T);
                    }
                    if (alive) {
                        emit_ray = true;
                        next_d = normalize(smp.wi);
                        next_o = add(h.position, scale(0.001f, h.normal));
                    }
                }
            }
            wb.state_t[pixel] = make_float4(T.x, T.y, T.z, __uint_as_float(rng));
            wb.state_c[pixel] = make_float4(C.x, C.y, C.z, 0.0f);
        }
    }
    uint32_t rs = wave_append(emit_ray, &wb.ext_count[k + 1]);
    uint32_t ss = wave_append(emit_shadow, &wb.sh_count[k]);
    if (emit_ray) {
        wb.ray_o[q ^ 1][rs] = make_float4(next_o.x, next_o.y, next_o.z, __uint_as_float(pixel));
        wb.ray_d[q ^ 1][rs] = make_float4(next_d.x, next_d.y, next_d.z, 0.0f);
    }
    if (emit_shadow) {
        wb.sh_o[ss] = make_float4(sh_o.x, sh_o.y, sh_o.z, __uint_as_float(pixel));
        wb.sh_d[ss] = make_float4(sh_d.x, sh_d.y, sh_d.z, sh_thr);
        wb.sh_c[ss] = make_float4(sh_c.x, sh_c.y, sh_c.z, 0.0f);
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
