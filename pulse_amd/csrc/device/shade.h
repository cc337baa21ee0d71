// shade.h -- one bounce of the path loop at a closest hit: the body of the
// reference's loop (path_tracer.wgsl:77-114 for path_trace, 173-196 for
// path_traceBBB) with the NEE shadow ray deferred to the shadow stage.
// Shared by the single-view and the batched stage kernels.
#pragma once

#include "pt_device.h"
#include "queue.h"

namespace pt {

struct ShadeOut {
    bool emit_ray = false, emit_shadow = false;
    V3 next_o{}, next_d{};          // next extension ray
    V3 sh_o{}, sh_d{}, sh_c{};      // shadow ray and the pixel colour if it is unoccluded
    float sh_thr = 0.0f;            // occlusion threshold (world t)
};

// Ray (o, d) hit {t, inst bits, tri bits, u} = ha, v = hv at trace index k of
// a path with throughput T, colour C and RNG state rng (all updated).
__device__ __forceinline__ ShadeOut shade_ray(const DevScene &sc, const FrameArgs &fa, int k, V3 o, V3 d, float4 ha,
                                              float hv, V3 &T, V3 &C, uint32_t &rng) {
    ShadeOut out;
    const float t = ha.x;
    HitInfo h = shade_hit(sc, o, d, t, __float_as_uint(ha.y), __float_as_uint(ha.z), ha.w, hv);
    Material m = load_material(sc.materials[h.material_index]);
    V3 wo = neg(d);
    const bool last = static_cast<uint32_t>(k) + 1u >= fa.max_traces;
    const bool lights = sc.light_count > 0u;
    if (fa.mode == PULSE_MODE_REF_NEE_LAMBERT) {
        if (k == 0 && !fa.gi) {
            // first hit stands in for the G-buffer sample
            // (path_tracer.wgsl:143-153)
            float a0 = rand_f(rng);
            float a1 = rand_f(rng);
            V3 scatter = sample_cosine_hemisphere(h.normal, a0, a1);
            V3 dl_vis = mk(0.0f, 0.0f, 0.0f);
            if (lights) {
                NeeSample ns = nee_sample(sc, h.position, h.normal, m, wo, false, rng, fa.light_power != 0u);
                out.emit_shadow = true;
                out.sh_o = ns.origin;
                out.sh_d = ns.dir;
                out.sh_thr = ns.threshold;
                dl_vis = ns.contrib;
            }
            out.sh_c = add(m.emissive, dl_vis);
            C = add(m.emissive, mk(0.0f, 0.0f, 0.0f));
            if (!lights) C = out.sh_c;
            T = m.base;
            if (!last) {
                out.emit_ray = true;
                out.next_o = add(h.position, scale(0.001f, h.normal));
                out.next_d = scatter;
            }
        } else {
            // loop body, path_tracer.wgsl:175-190
            if (lights) {
                NeeSample ns = nee_sample(sc, h.position, h.normal, m, wo, false, rng, fa.light_power != 0u);
                out.emit_shadow = true;
                out.sh_o = ns.origin;
                out.sh_d = ns.dir;
                out.sh_thr = ns.threshold;
                out.sh_c = add(C, mul(T, ns.contrib));
            }
            // without lights the reference adds throughput * 0 too
            C = add(C, mul(T, mk(0.0f, 0.0f, 0.0f)));
            // the GI pass (gi.wgsl:112-120) draws roulette and scatter at the
            // last depth too: its RNG runs on into the next sample
            if (!last || fa.gi) {
                T = mul(T, m.base);
                bool alive = true;
                if (fa.gi || static_cast<uint32_t>(k - 1) >= fa.rr_start) {
                    float p = max3(T);
                    if (rand_f(rng) > p) alive = false;
                    else T = scale(1.0f / p, T);
                }
                if (alive) {
                    float b0 = rand_f(rng);
                    float b1 = rand_f(rng);
                    out.emit_ray = !last;
                    out.next_d = sample_cosine_hemisphere(h.normal, b0, b1);
                    out.next_o = add(h.position, scale(0.001f, h.normal));
                }
            }
        }
    } else {
        if (fa.mode == PULSE_MODE_REF_GGX || k == 0) C = add(C, mul(T, m.emissive));
        if (fa.mode == PULSE_MODE_GGX_NEE && lights) {
            NeeSample ns = nee_sample(sc, h.position, h.normal, m, wo, fa.nee_ggx != 0u, rng, fa.light_power != 0u);
            out.emit_shadow = true;
            out.sh_o = ns.origin;
            out.sh_d = ns.dir;
            out.sh_thr = ns.threshold;
            out.sh_c = add(C, mul(T, ns.contrib));
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
                out.emit_ray = true;
                out.next_d = normalize(smp.wi);
                out.next_o = add(h.position, scale(0.001f, h.normal));
            }
        }
    }
    return out;
}

}  // namespace pt
