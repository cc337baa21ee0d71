// adjoint.h -- replay adjoint of the L2 image loss with respect to per-
// material albedo (base_color.rgb), GGX roughness and emissive.rgb.
//
// The paths are those of PULSE_MODE_GGX_NEE (emission at the first hit,
// light sampling at every hit), of the reference's own path_trace,
// PULSE_MODE_REF_GGX (path_tracer.wgsl:58-124: emission at every hit, no
// light sampling), or of its path_traceBBB, PULSE_MODE_REF_NEE_LAMBERT
// (path_tracer.wgsl:128-202: cosine sampling, Lambert light sampling,
// emission at the first hit; no roughness dependence).
//
// Each path is replayed from its seed (the pixel index and accumulation
// count, path_tracer.wgsl:60-61), so it takes exactly the forward
// decisions.  Its closest hits and shadow-ray visibilities come from the
// forward pass's record when it kept one (PULSE_FLAG_RECORD_PATHS, 21 bytes
// per vertex: no traversal at all), else from the same f32 traversal run
// again.  Along
// the replay the vertex computations run in Dual<MAXV> arithmetic (MAXV the
// kernel's vertex capacity): the tangent k is the derivative with respect to the roughness
// used at vertex k, which moves the sampled half vector and therefore every
// later ray, hit point, normal, NEE term and BSDF weight.  Albedo and emission
// enter linearly; they are handled by a reverse sweep over the stored vertex
// weights: with C = sum_k beta_k c_k, beta_{k+1} = beta_k w_k and
// S_k = c_k + w_k S_{k+1}, dC/dc_j = beta_j and dC/dw_j = beta_j S_{j+1}.
//
// The derivative is that of the estimator with every decision held fixed
// (closest-hit ids, lobe, occlusion, fold of the light sample), i.e. the
// derivative a finite difference with the same random numbers measures away
// from those discontinuities.  Russian roulette (path_tracer.wgsl:104-108)
// is replayed with its decision.  Its p = max(throughput) divides the
// throughput: where p > 1 the path survives whatever the draw and 1/p is
// differentiated like any weight; where p <= 1 the survival is random and
// its score d ln p is added, which cancels 1/p's own derivative in every
// later contribution (E[1{u < p} / p] = 1) but not in later roulettes: a
// later p > 1 is measured on the throughput divided by this p, so its
// 1/p' carries this p's derivative back.
//
// The light picked for next event estimation is a decision too; its
// probability (uniform, or by emission strength with
// PULSE_FLAG_LIGHT_POWER) divides the light sample's contribution, which
// ends there: E over the pick of contribution / P_pick does not depend on
// P_pick, so the pdf is held at its value.
//
// The GGX lobe choice is a decision of the same kind (probability P = F0max
// or 1 - F0max, which follows the albedo of a metallic material and, with
// moving directions, the roughness): the weight holds P fixed, the terms
// that do not cancel come in like those of a random survival.
//
// Reverse sweep: with a_k = dL/d ln(beta_k) per channel (beta_k the
// throughput reaching vertex k) and S_k the suffix sum of vertex
// contributions, a_k = g beta_k S_k + E_k, where E carries the roulette
// terms: E_k = E_{k+1} - e_c X_k for a survived roulette at k (c the
// channel of the maximum), X_k = sum over channels of a_{k+1} where p > 1
// and of E_{k+1} where p <= 1 (the score removes g beta_{k+1} S_{k+1});
// a parameter moving ln w_k (the weight of vertex k) by d then adds
// (g beta_{k+1} S_{k+1} + E_k) . d.
//
// Per block the gradient is accumulated in LDS (float atomics), then added to
// the global gradient with one atomic per entry; a gradient too large for the
// LDS is added to device memory directly.  Either way each add first sums the
// lanes of a wave that hold the same material (grad_add_wave).
#pragma once

#include <type_traits>

#include "queue.h"
#include "traverse.h"

namespace pt {

constexpr int kAdjMaxVertices = 4;      // path vertices of the default kernels
constexpr int kAdjMaxVerticesDeep = 8;  // the deep kernels (fixed-direction roughness only)
constexpr int kGradPerMaterial = 7;     // base.rgb, roughness, emissive.rgb

// One launch covers up to kAdjMaxViews views of the bound scene; the loss is
// the sum of the views' losses, so the gradient is the sum of theirs.  Block
// b of view v replays the view's pixels [128 (b - block0), 128 (b - block0 +
// 1)) in row order (measured faster than 8x8 squares per wave; a wave's 64
// neighbouring primary rays still mostly hit the same instances, which the
// wave sums of the gradient adds exploit).
constexpr int kAdjMaxViews = 16;
constexpr uint32_t kAdjBlockPixels = 128;

struct AdjView {
    pulse_camera cam;
    const float4 *image;   // forward image (running mean over spp samples)
    const float4 *target;  // target image
    float4 *replay;        // optional: the replayed running mean (verification)
    float grad_scale;      // dL/dI = grad_scale * (I - T): 2 / (3 npix) of this view
    uint32_t npix;         // width * height
    uint32_t block0;       // the view's first block
    // the forward pass's recorded paths (null: trace them again): at
    // (sample * npix + pixel) * vertices + vertex, the closest hit
    // {t, inst bits, tri bits, u}, its v, and the shadow ray's visibility
    const float4 *rec_hit;
    const float *rec_v;
    const uint8_t *rec_vis;
};

struct AdjArgs {
    AdjView views[kAdjMaxViews];
    uint32_t n_views;
    uint32_t vertices;     // path vertices: max_depth (GGX modes) or max_depth + 1 (Lambert: the first hit and max_depth bounces)
    uint32_t nee_ggx;
    uint32_t spp;          // samples replayed per pixel: 0 .. spp - 1 (accumulation_count + the call's spp)
    uint32_t light_power;  // lights picked by emission strength (its pdf is held: see below)
    uint32_t ref_ggx;      // PULSE_MODE_REF_GGX: emission at every hit, no light sampling
    uint32_t lambert;      // PULSE_MODE_REF_NEE_LAMBERT: cosine sampling, weight = albedo

    float *grad;           // n_materials * 7, accumulated into
    uint32_t n_materials;
    uint32_t use_lds;      // gradient buffer fits in LDS
    uint32_t horizon_clamp;
    uint32_t detached;     // roughness derivative at fixed sampled directions
    uint32_t reduce_rounds;  // distinct keys per wave summed before the atomic (0: per-lane atomics)
    uint32_t rr_start;     // Russian roulette from this depth on (as the forward, shade.h)
};

template <class S>
__device__ __forceinline__ V3 val3(const V3T<S> &a) { return mk(val(a.x), val(a.y), val(a.z)); }

// The intersection of the traversal (traverse.h / pt_device.h: object-space
// ray, Moeller-Trumbore, world t through the hit point) recomputed for a
// known triangle in scalar type S.  Values equal the traversal's bit for bit.
template <class S>
__device__ __forceinline__ void intersect_known(const DevScene &sc, const V3T<S> &o, const V3T<S> &d, uint32_t inst,
                                                uint32_t tri, S &tw, S &u, S &v) {
    const pulse_mesh_instance &mi = sc.instances[inst];
    float wm[16], ow[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        wm[i] = mi.world_object[i];
        ow[i] = mi.object_world[i];
    }
    V3T<S> oo = transform_position(wm, o);
    V3T<S> dd = transform_direction(wm, d);
    S dlen = length(dd);
    V3T<S> od = divs(dd, dlen);
    uint32_t slot = mi.index_offset + tri;
    float4 A = sc.tri_bvh[3 * slot], B = sc.tri_bvh[3 * slot + 1], C = sc.tri_bvh[3 * slot + 2];
    V3 p0 = mk(A.x, A.y, A.z), e1 = mk(A.w, B.x, B.y), e2 = mk(B.z, B.w, C.x);
    V3T<S> hv = cross(od, e2);
    S a = dot(e1, hv);
    S f = 1.0f / a;
    V3T<S> s = sub(oo, p0);
    u = f * dot(s, hv);
    V3T<S> q = cross(s, e1);
    v = f * dot(od, q);
    S t = f * dot(e2, q);
    V3T<S> hit_world = transform_position(ow, add(oo, scale(t, od)));
    tw = length(sub(hit_world, o));
}

constexpr uint32_t kNoKey = 0xffffffffu;

// Sum of x over the lanes with `on` set, in every lane (butterfly over the
// 64 lanes; every lane of the wave must be active).
__device__ __forceinline__ float wave_sum(float x, bool on) {
    float s = on ? x : 0.0f;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    return s;
}

// Adds v[j] to buf[key + j], j < N, for every lane whose key is not kNoKey.
// First up to `rounds` distinct keys of the wave (in lane order: the lowest
// pending lane's key each round) are summed over the lanes holding them and
// added by one atomic per entry; the lanes still pending then add their own
// values.  Floating-point sums over lanes in another order than per-lane
// atomics would take: the result equals those up to rounding.  Every lane of
// the wave must reach the call (the kernels call it from uniform control
// flow; dead lanes pass kNoKey).
template <int N>
__device__ __forceinline__ void grad_add_wave(float *buf, uint32_t key, const float (&v)[N], uint32_t rounds) {
    const uint32_t lane = __lane_id();
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint64_t pend = __ballot(key != kNoKey);
        if (pend == 0) return;
        const int leader = __builtin_ctzll(pend);
        const uint32_t k = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(key), leader));
        const bool mine = key == k;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float s = wave_sum(v[j], mine);
            if (lane == static_cast<uint32_t>(leader) && s != 0.0f) atomicAdd(buf + k + j, s);
        }
        if (mine) key = kNoKey;
    }
    if (key != kNoKey) {
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (v[j] != 0.0f) atomicAdd(buf + key + j, v[j]);
    }
}

// DETACHED = true: the roughness derivative holds every sampled direction
// fixed and differentiates the sampling weight brdf(wi; r) / pdf with the pdf
// detached, i.e. w * d ln(D G)/dr at each specular vertex (and at GGX-evaluated
// NEE samples).  Unbiased for the derivative of the expected image including
// the effect of visibility changes, which the reparameterised form (DETACHED =
// false, tangents through the moving directions) does not capture.
// LOOP: the traversal of the forward pass being replayed (1: the binary
// BVHs, 2: the 4-wide ones), so every replayed path is the forward's.
// MAXV: path vertices the per-vertex arrays hold (a.vertices at most MAXV).
template <int BLOCK, int STACK, bool DETACHED, int LOOP = 1, int MAXV = kAdjMaxVertices>
__global__ void __launch_bounds__(BLOCK) k_adjoint(DevScene sc, AdjArgs a) {
    static_assert(BLOCK == kAdjBlockPixels, "a block replays 128 pixels");
    static_assert(MAXV <= 8, "the roulette record holds 4 bits per vertex in 32");
    using S = typename std::conditional<DETACHED, float, Dual<MAXV>>::type;
    __shared__ uint32_t smem[STACK * BLOCK];
    extern __shared__ float sgrad[];
    const uint32_t nent = a.n_materials * kGradPerMaterial;
    if (a.use_lds) {
        for (uint32_t i = threadIdx.x; i < nent; i += BLOCK) sgrad[i] = 0.0f;
        __syncthreads();
    }
    float *gbuf = a.use_lds ? sgrad : a.grad;
    // the block's view and pixels; lanes past the view's last pixel run the
    // sample loop without vertices (the gradient adds need the whole wave)
    uint32_t vi = 0;
    for (uint32_t i = 1; i < a.n_views; ++i)
        if (blockIdx.x >= a.views[i].block0) vi = i;
    const AdjView &vw = a.views[vi];
    const uint32_t p = (blockIdx.x - vw.block0) * kAdjBlockPixels + threadIdx.x;
    const uint32_t W = vw.cam.width;
    const bool live = p < vw.npix;
    const uint32_t pixel = live ? p : 0u;
    const uint32_t x = pixel % W, y = pixel / W;
    const uint32_t depth_end = live ? a.vertices : 0u;
    const uint32_t rounds = a.reduce_rounds;
    {
        LdsStack stk{smem + threadIdx.x, BLOCK, STACK};
        float4 im = make_float4(0.0f, 0.0f, 0.0f, 0.0f), tg = im;
        if (live) {
            im = vw.image[pixel];
            tg = vw.target[pixel];
        }
        // dL/dC of one sample: dL/dI * dI/dC, I the mean of spp samples
        const float inv_spp = 1.0f / static_cast<float>(a.spp);
        const float gx = vw.grad_scale * (im.x - tg.x) * inv_spp;
        const float gy = vw.grad_scale * (im.y - tg.y) * inv_spp;
        const float gz = vw.grad_scale * (im.z - tg.z) * inv_spp;
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        for (uint32_t s = 0; s < a.spp; ++s) {
            uint32_t rng = pixel_seed(pixel, s);
            V3 o0, d0;
            primary_ray(vw.cam, x, y, o0, d0);
            V3T<S> o = lift<S>(o0), d = lift<S>(d0);
            V3T<S> T = lift<S>(mk(1.0f, 1.0f, 1.0f));
            V3T<S> C = lift<S>(mk(0.0f, 0.0f, 0.0f));
            // stored vertex values for the reverse sweep
            V3 beta[MAXV], cvert[MAXV], wvert[MAXV];
            V3 dDdb[MAXV], dDdLe[MAXV], dDdr[MAXV];
            float dwdb[MAXV], dlogw[MAXV], dlogq[MAXV], dlogp[MAXV];
            // 3 bits per vertex: the channel of F0max at the lobe choice | 4 if
            // the pdf floor replaced the density
            uint32_t lobe = 0;
            uint32_t mat[MAXV], lmat[MAXV];
            bool vis[MAXV];
            // 4 bits per vertex: 1 + the channel of the maximum at a survived
            // roulette (0: none), | 4 if its p exceeded 1
            uint32_t rrc = 0;
            // roughness tangent of ln(throughput with every 1/p varying) -
            // ln(T) (one scalar: the p <= 1 survivals hold T's 1/p fixed)
            S nrm = sconst<S>(0.0f);
            int nv = 0;
            for (uint32_t depth = 0; depth < depth_end; ++depth) {
                float t = kTFar, u = 0.0f, v = 0.0f;
                uint32_t inst = 0, tri = 0;
                const uint32_t rec = (s * vw.npix + pixel) * a.vertices + depth;
                if (vw.rec_hit) {
                    const float4 hr = vw.rec_hit[rec];
                    t = hr.x;
                    inst = __float_as_uint(hr.y);
                    tri = __float_as_uint(hr.z);
                    u = hr.w;
                    v = vw.rec_v[rec];
                } else {
                    trace_any<LOOP, false>(sc, val3(o), val3(d), t, inst, tri, u, v, 0.0f, stk);
                }
                if (!(t < kTFar)) break;
                S tw, uu, vv;
                if constexpr (DETACHED) {
                    tw = t;
                    uu = u;
                    vv = v;
                } else {
                    intersect_known(sc, o, d, inst, tri, tw, uu, vv);
                }
                HitInfoT<S> h = shade_hit(sc, o, d, tw, inst, tri, uu, vv);
                const pulse_material &pm = sc.materials[h.material_index];
                MaterialT<S> m;
                m.base = ld3(pm.base_color);
                m.emissive = ld3(pm.emissive);
                if constexpr (DETACHED) {
                    m.roughness = pm.perceptual_roughness;
                } else {
                    m.roughness = dvar<MAXV>(pm.perceptual_roughness, static_cast<int>(depth));
                }
                m.reflectance = pm.reflectance;
                m.metallic = pm.metallic;
                V3T<S> wo = neg(d);
                V3 ck = mk(0.0f, 0.0f, 0.0f);
                beta[nv] = val3(T);
                mat[nv] = h.material_index;
                if (depth == 0 || a.ref_ggx) {
                    C = add(C, mul(T, m.emissive));
                    ck = m.emissive;
                }
                vis[nv] = false;
                dDdb[nv] = mk(0.0f, 0.0f, 0.0f);
                dDdLe[nv] = mk(0.0f, 0.0f, 0.0f);
                dDdr[nv] = mk(0.0f, 0.0f, 0.0f);
                lmat[nv] = 0;
                float pre0 = 0.0f, pre1 = 0.0f;
                if (a.lambert && depth == 0) {
                    // path_traceBBB draws the first hit's cosine sample before its light sample
                    pre0 = rand_f(rng);
                    pre1 = rand_f(rng);
                }
                if (sc.light_count > 0u && !a.ref_ggx) {
                    NeeSampleT<S> ns = nee_sample(sc, h.position, h.normal, m, wo, a.nee_ggx != 0u, rng, a.light_power != 0u);
                    float tt = kTFar, su = 0.0f, sv = 0.0f;
                    uint32_t si = 0, st = 0;
                    const bool occ = vw.rec_vis ? vw.rec_vis[rec] == 0u
                                               : trace_any<LOOP, true>(sc, val3(ns.origin), val3(ns.dir), tt, si, st,
                                                                       su, sv, val(ns.threshold), stk);
                    if (!occ) {
                        C = add(C, mul(T, ns.contrib));
                        ck = add(ck, val3(ns.contrib));
                        vis[nv] = true;
                        dDdb[nv] = ns.d_base;
                        dDdLe[nv] = ns.d_emissive;
                        dDdr[nv] = ns.d_rough;
                        lmat[nv] = ns.light_material;
                    } else {
                        C = add(C, mul(T, mk(0.0f, 0.0f, 0.0f)));
                    }
                }
                cvert[nv] = ck;
                wvert[nv] = mk(0.0f, 0.0f, 0.0f);
                dwdb[nv] = 0.0f;
                dlogw[nv] = 0.0f;
                dlogq[nv] = 0.0f;
                dlogp[nv] = 0.0f;
                const bool last = depth + 1u >= a.vertices;
                if (!last) {
                    V3T<S> next_d;
                    bool roulette;
                    if (a.lambert) {
                        // path_traceBBB (path_tracer.wgsl:128-202): the
                        // cosine-sampled Lambert weight is the albedo; the
                        // roulette runs from loop depth rr_start (vertex
                        // rr_start + 1), never at the first hit
                        T = mul(T, lift<S>(m.base));
                        wvert[nv] = m.base;
                        dwdb[nv] = 1.0f;
                        roulette = depth >= 1u && depth - 1u >= a.rr_start;
                    } else {
                        SampleT<S> smp = importance_sample_ggx_d(h.normal, wo, m, rng);
                        if (a.horizon_clamp && below_horizon(smp)) {
                            nv += 1;
                            break;
                        }
                        T = mul(T, smp.reflectance);
                        wvert[nv] = val3(smp.reflectance);
                        dwdb[nv] = smp.dbase;
                        dlogw[nv] = smp.dlogf_dr;
                        dlogq[nv] = smp.dlogq_dr;
                        dlogp[nv] = smp.dlogp_db;
                        lobe |= (smp.lobe_c0 | (smp.floored ? 4u : 0u)) << (3 * nv);
                        if constexpr (!DETACHED) {
                            // the weight holds the lobe probability P fixed: the
                            // throughput with every factor varying has d ln P less
                            if (val(smp.lobe_p) > 0.0f) nrm = nrm + (1.0f - smp.lobe_p / val(smp.lobe_p));
                        }
                        next_d = normalize(smp.wi);
                        roulette = depth >= a.rr_start;
                    }
                    if (roulette) {
                        // the forward's roulette (shade.h) on the replayed
                        // throughput: the same draw, the same end of the
                        // path, the same 1/p
                        const V3 tv = val3(T);
                        const float pr = max3(tv);
                        if (rand_f(rng) > pr) {
                            nv += 1;
                            break;
                        }
                        const float ip = 1.0f / pr;
                        const uint32_t c = tv.x == pr ? 0u : (tv.y == pr ? 1u : 2u);
                        rrc |= ((c + 1u) | (pr > 1.0f ? 4u : 0u)) << (4 * nv);
                        if constexpr (DETACHED) {
                            T = scale(ip, T);
                        } else {
                            const S tc = c == 0u ? T.x : (c == 1u ? T.y : T.z);
                            if (pr > 1.0f) {
                                // 1/p with p = T_c times the normaliser
                                // (value ip, as the forward)
                                const S ipd = 1.0f / (tc * (1.0f + nrm));
                                T = scale(ipd, T);
                            } else {
                                nrm = 1.0f - tc / pr;
                                T = scale(ip, T);
                            }
                        }
                        wvert[nv] = scale(ip, wvert[nv]);
                        dwdb[nv] *= ip;
                    }
                    if (a.lambert) {
                        const float b0 = depth == 0 ? pre0 : rand_f(rng);
                        const float b1 = depth == 0 ? pre1 : rand_f(rng);
                        next_d = sample_cosine_hemisphere(h.normal, b0, b1);
                    }
                    d = next_d;
                    o = add(h.position, scale(0.001f, h.normal));
                }
                nv += 1;
            }
            // Both sweeps visit the vertex slots the wave's longest path has,
            // in every lane, so the wave-wide adds see the whole wave; a lane
            // whose path is shorter reads nothing and adds nothing there.
            int kmax = 0;
#pragma unroll
            for (int k = 0; k < MAXV; ++k)
                if (__ballot(nv > k) != 0) kmax = k + 1;
            if constexpr (!DETACHED) {
                // roughness: tangents of the replayed colour
                for (int k = 0; k < kmax; ++k) {
                    uint32_t key = kNoKey;
                    float v1[1] = {0.0f};
                    if (k < nv) {
                        const float dr = gx * C.x.d[k] + gy * C.y.d[k] + gz * C.z.d[k];
                        if (dr == dr) {
                            key = mat[k] * kGradPerMaterial + 3;
                            v1[0] = dr;
                        }
                    }
                    grad_add_wave(gbuf, key, v1, rounds);
                }
            }
            // albedo and emission: reverse sweep over the vertices
            float Sx = 0.0f, Sy = 0.0f, Sz = 0.0f;
            float Ex = 0.0f, Ey = 0.0f, Ez = 0.0f;  // E_{k+1}: the 1/p terms (zero without p > 1)
            V3 bn = mk(0.0f, 0.0f, 0.0f);           // beta_{k+1}
            for (int k = kmax - 1; k >= 0; --k) {
                uint32_t kb = kNoKey, ke = kNoKey, k0 = kNoKey;  // albedo+roughness, NEE emission, direct emission
                float v4[4] = {0.0f, 0.0f, 0.0f, 0.0f}, e3[3] = {0.0f, 0.0f, 0.0f}, d3[3] = {0.0f, 0.0f, 0.0f};
                if (k < nv) {
                    const V3 b = beta[k];
                    const uint32_t m = mat[k];
                    const float dwb = dwdb[k];
                    const V3 wk = wvert[k];
                    // F = E_k (the roulette terms of this vertex and later ones)
                    float Fx = Ex, Fy = Ey, Fz = Ez;
                    const uint32_t ck = (rrc >> (4 * k)) & 7u;
                    if (ck != 0u) {
                        const float es = Ex + Ey + Ez;
                        const float as = (ck & 4u) ? gx * bn.x * Sx + gy * bn.y * Sy + gz * bn.z * Sz + es : es;
                        Fx -= (ck & 3u) == 1u ? as : 0.0f;
                        Fy -= (ck & 3u) == 2u ? as : 0.0f;
                        Fz -= (ck & 3u) == 3u ? as : 0.0f;
                    }
                    // d ln w_k / d base = dwb / w_k per channel; F is nonzero only
                    // in channels whose w_k is positive (they hold a later p > 1)
                    float ax = gx * b.x * (dDdb[k].x + Sx * dwb) + (Fx != 0.0f ? Fx * dwb / wk.x : 0.0f);
                    float ay = gy * b.y * (dDdb[k].y + Sy * dwb) + (Fy != 0.0f ? Fy * dwb / wk.y : 0.0f);
                    float az = gz * b.z * (dDdb[k].z + Sz * dwb) + (Fz != 0.0f ? Fz * dwb / wk.z : 0.0f);
                    if (dlogp[k] != 0.0f) {
                        // the lobe probability (metallic: it follows base_c0): -d ln P
                        // in the weight reaches the 1/p terms only; where the
                        // pdf floor left P out of the weight, the choice's score
                        // reaches every later contribution
                        const uint32_t lb = (lobe >> (3 * k)) & 7u;
                        const float t = (lb & 4u) ? (gx * bn.x * Sx + gy * bn.y * Sy + gz * bn.z * Sz) * dlogp[k]
                                                  : -(Fx + Fy + Fz) * dlogp[k];
                        ax += (lb & 3u) == 0u ? t : 0.0f;
                        ay += (lb & 3u) == 1u ? t : 0.0f;
                        az += (lb & 3u) == 2u ? t : 0.0f;
                    }
                    if (ax == ax && ay == ay && az == az) {
                        kb = m * kGradPerMaterial;
                        v4[0] = ax;
                        v4[1] = ay;
                        v4[2] = az;
                    }
                    if constexpr (DETACHED) {
                        // roughness at fixed directions: NEE (GGX evaluator) and
                        // the sampling weight w_k with the likelihood ratio of
                        // its direction (d (w q) = w q d ln f, every channel);
                        // the 1/p terms follow w_k alone (p is the path's own
                        // throughput), d ln w = d ln f - d ln q
                        const float dl = dlogw[k];
                        const float dr = gx * b.x * (dDdr[k].x + Sx * wk.x * dl) +
                                         gy * b.y * (dDdr[k].y + Sy * wk.y * dl) +
                                         gz * b.z * (dDdr[k].z + Sz * wk.z * dl) +
                                         (Fx + Fy + Fz) * (dl - dlogq[k]);
                        if (dr == dr) {
                            kb = m * kGradPerMaterial;
                            v4[3] = dr;
                        }
                    }
                    if (vis[k]) {
                        const float ex = gx * b.x * dDdLe[k].x, ey = gy * b.y * dDdLe[k].y, ez = gz * b.z * dDdLe[k].z;
                        if (ex == ex && ey == ey && ez == ez) {
                            ke = lmat[k] * kGradPerMaterial + 4;
                            e3[0] = ex;
                            e3[1] = ey;
                            e3[2] = ez;
                        }
                    }
                    if (k == 0 || a.ref_ggx) {
                        // emission seen directly (the first hit; every hit in
                        // the reference's path_trace): dC/dLe = the throughput
                        const float ex = gx * b.x, ey = gy * b.y, ez = gz * b.z;
                        if (ex == ex && ey == ey && ez == ez) {
                            k0 = m * kGradPerMaterial + 4;
                            d3[0] = ex;
                            d3[1] = ey;
                            d3[2] = ez;
                        }
                    }
                    const V3 cv = cvert[k];
                    Sx = cv.x + wk.x * Sx;
                    Sy = cv.y + wk.y * Sy;
                    Sz = cv.z + wk.z * Sz;
                    Ex = Fx;
                    Ey = Fy;
                    Ez = Fz;
                    bn = b;
                }
                grad_add_wave(gbuf, kb, v4, rounds);
                grad_add_wave(gbuf, ke, e3, rounds);
                if (k == 0 || a.ref_ggx) grad_add_wave(gbuf, k0, d3, rounds);
            }
            if (vw.replay) {
                // running mean of path_tracer.wgsl:116-123 over the replayed samples
                float w = 1.0f / (static_cast<float>(s) + 1.0f);
                float omw = 1.0f - w;
                float ox = s > 0 ? acc.x : 0.0f, oy = s > 0 ? acc.y : 0.0f, oz = s > 0 ? acc.z : 0.0f;
                acc = make_float4(ox * omw + val(C.x) * w, oy * omw + val(C.y) * w, oz * omw + val(C.z) * w, 1.0f);
            }
        }
        if (vw.replay && live) vw.replay[pixel] = acc;
    }
    if (a.use_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nent; i += BLOCK)
            if (sgrad[i] != 0.0f) atomicAdd(a.grad + i, sgrad[i]);
    }
}

// Loss L = mean over pixels and rgb of (I - T)^2, accumulated in f64.
__global__ void __launch_bounds__(256) k_image_loss(const float4 *image, const float4 *target, uint32_t npix,
                                                    double *out) {
    __shared__ double part[256];
    double s = 0.0;
    for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < npix; p += gridDim.x * 256) {
        float4 i = image[p], t = target[p];
        double dx = static_cast<double>(i.x) - t.x, dy = static_cast<double>(i.y) - t.y,
               dz = static_cast<double>(i.z) - t.z;
        s += dx * dx + dy * dy + dz * dz;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0] / (3.0 * static_cast<double>(npix)));
}

// Zero-fill (used instead of a memset so the sequence can be captured in a
// graph).
__global__ void k_fill_zero(float *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace pt
