// traverse.h -- single-loop two-level traversal for wave64.
//
// Same visiting order and results as traverse_tlas/traverse_blas (the
// reference's nested loops, utilities.wgsl:127-188, 259-326), restructured so
// that every lane of a wave executes the same "test the two children of the
// current node" step whether it is in the TLAS or inside an instance's BLAS:
// the current level's ray, node array and culling bound live in the same
// registers, and a level change only swaps them.  The nested form leaves the
// lanes that are still in the TLAS idle while others walk a BLAS.
//
// Division: the slab test divides by the ray direction 6 times per box.  With
// a correctly rounded reciprocal r = 1/d, q0 = x*r, e = fma(-q0, d, x),
// q = fma(e, r, q0) equals x/d rounded to nearest (Markstein's theorem) as long
// as nothing underflows or overflows.  That is guaranteed when every node
// coordinate and origin component is 0 or has magnitude in [2^-56, 2^40] and
// every direction component has magnitude in [2^-40, 2^40] (then |x/d| lies in
// [2^-120, 2^81]); rays or scenes outside that range use true division.
#pragma once

#include "pt_device.h"

namespace pt {

struct PreRay {
    V3 o, d, r;  // origin, direction, correctly rounded 1/d
    bool fast;   // reciprocal-multiply division is exact for this ray
};

__device__ __forceinline__ bool in_range_or_zero(float v, float lo, float hi) {
    float a = fabsf(v);
    return v == 0.0f || (a >= lo && a <= hi);
}
__device__ __forceinline__ bool in_range(float v, float lo, float hi) {
    float a = fabsf(v);
    return a >= lo && a <= hi;
}

__device__ __forceinline__ PreRay prep_ray(V3 o, V3 d, bool scene_fast) {
    PreRay p;
    p.o = o;
    p.d = d;
    p.r = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    p.fast = scene_fast && in_range(d.x, 0x1p-40f, 0x1p40f) && in_range(d.y, 0x1p-40f, 0x1p40f) &&
             in_range(d.z, 0x1p-40f, 0x1p40f) && in_range_or_zero(o.x, 0x1p-56f, 0x1p40f) &&
             in_range_or_zero(o.y, 0x1p-56f, 0x1p40f) && in_range_or_zero(o.z, 0x1p-56f, 0x1p40f);
    return p;
}

__device__ __forceinline__ float qdiv(float x, float d, float r) {
    float q0 = x * r;
    float e = __builtin_fmaf(-q0, d, x);
    return __builtin_fmaf(e, r, q0);
}

// ray_aabb_intersect (utilities.wgsl:405-426) with the exact fast division.
__device__ __forceinline__ float slab_pre(const PreRay &p, float rec_t, float4 lo, float4 hi) {
    if (!p.fast) return slab(p.o, p.d, rec_t, lo, hi);
    float tx1 = qdiv(lo.x - p.o.x, p.d.x, p.r.x);
    float tx2 = qdiv(hi.x - p.o.x, p.d.x, p.r.x);
    float tmin = fminf(tx1, tx2);
    float tmax = fmaxf(tx1, tx2);
    float ty1 = qdiv(lo.y - p.o.y, p.d.y, p.r.y);
    float ty2 = qdiv(hi.y - p.o.y, p.d.y, p.r.y);
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = qdiv(lo.z - p.o.z, p.d.z, p.r.z);
    float tz2 = qdiv(hi.z - p.o.z, p.d.z, p.r.z);
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    return (tmax >= tmin && tmin < rec_t && tmax > 0.0f) ? tmin : kTFar;
}

// Resumable form of the single-loop traversal: begin() sets a ray up, each
// step() visits one node (or tests TRIS_PER_STEP triangles of a leaf, in
// order, or enters/leaves one instance), so a persistent kernel can hand a
// lane a new ray as soon as its ray is done.
//
// Closest hit (ANY_HIT = false): rec_* hold the closest hit so far.
// Occlusion (ANY_HIT = true): `occluded` becomes true iff the reference's
// closest-hit t would be < threshold.  Two exact shortcuts are taken in
// occlusion mode: the TLAS and each BLAS are culled at a padded threshold
// (nothing culled can convert to a world t below the threshold), and a BLAS
// stops at the first hit whose object t is below a bound that converts to a
// world t below the threshold for certain (the instance's closest hit is
// then below it too).
template <bool ANY_HIT, int TRIS_PER_STEP = 1>
struct Traversal {
    V3 wo, wd;
    PreRay world;
    float threshold, slack_thr, world_cull;
    PreRay cur;        // ray of the current level
    float cur_t;       // culling bound of the current level
    const float4 *nodes;
    bool in_blas;
    int sp, sp_base, iter_t, iter_b;
    uint32_t leaf_next, leaf_end, inst, idx_off;
    uint32_t tri_next, tri_end;  // triangles of the BLAS leaf in progress
    bool want_advance;           // next: leave the BLAS and/or enter the next instance
    float dlen, start_t, safe_t;
    ObjHit h;
    uint32_t node_first, node_count;  // current node: child pair / first primitive, primitive count
    float rec_t, rec_u, rec_v;
    uint32_t rec_inst, rec_tri;
    bool occluded;

    // false: nothing to traverse (empty TLAS); the result is already final.
    __device__ __forceinline__ bool begin(const DevScene &sc, V3 o, V3 d, float t_max, float thr) {
        rec_t = t_max;
        rec_inst = 0;
        rec_tri = 0;
        rec_u = 0.0f;
        rec_v = 0.0f;
        occluded = false;
        load_node(sc.tlas_nodes, 0u);
        if (node_first == 0u && node_count == 0u) return false;
        wo = o;
        wd = d;
        world = prep_ray(wo, wd, sc.fast_div);
        threshold = thr;
        world_cull = rec_t;
        slack_thr = 0.0f;
        if (ANY_HIT) {
            slack_thr = world_bound_slack(wo, threshold);
            world_cull = threshold + slack_thr;
        }
        cur = world;
        cur_t = world_cull;
        nodes = sc.tlas_nodes;
        in_blas = false;
        sp = 0;
        sp_base = 0;
        iter_t = 0;
        iter_b = 0;
        leaf_next = 0;
        leaf_end = 0;
        tri_next = 0;
        tri_end = 0;
        want_advance = false;
        inst = 0;
        idx_off = 0;
        dlen = 1.0f;
        start_t = kTFar;
        safe_t = -1.0f;
        h.t = kTFar;
        h.found = false;
        h.tri = 0;
        h.u = 0.0f;
        h.v = 0.0f;
        return true;
    }

    // One step; false when the traversal is complete.  A lane whose next
    // operation is an instance change (leave a finished BLAS, enter the next
    // instance of a TLAS leaf) waits while run_advance is false: the caller
    // may gather such lanes so that the wave runs that long code path for
    // many lanes at once.  Waiting does not change the ray's own sequence of
    // operations.
    __device__ __forceinline__ bool step(const DevScene &sc, LdsStack &stk, bool run_advance = true) {
        if (want_advance) {
            if (!run_advance) return true;
            want_advance = false;
            return advance(sc, stk);
        }
        const uint32_t first = node_first;
        const uint32_t count = node_count;
        bool advance_instance = false;  // BLAS finished: convert and move on
        bool do_pop = false;
        if (!in_blas) {
            if (iter_t >= kMaxIterations) {
                atomicOr(sc.error_flags, kErrIterationCap);
                return false;
            }
            iter_t += 1;
            if (count > 0u) {
                leaf_next = first;
                leaf_end = first + count;
                advance_instance = true;  // enter leaf_next (no BLAS finished yet)
                inst = 0xffffffffu;
            }
        } else {
            // BLAS: a leaf's triangles are tested one per step, in order, so
            // lanes at leaves and lanes at interior nodes share each step's cost
            bool test_tri = tri_next < tri_end;
            if (!test_tri) {
                if (iter_b >= kMaxIterations) {
                    atomicOr(sc.error_flags, kErrIterationCap);
                    advance_instance = true;
                } else {
                    iter_b += 1;
                    if (count > 0u) {
                        tri_next = first;
                        tri_end = first + count;
                        test_tri = true;
                    }
                }
            }
            if (test_tri) {
                mt(cur.o, cur.d, sc.tri_bvh, idx_off + tri_next, tri_next, h);
                tri_next += 1;
                if (TRIS_PER_STEP == 2 && tri_next < tri_end) {
                    mt(cur.o, cur.d, sc.tri_bvh, idx_off + tri_next, tri_next, h);
                    tri_next += 1;
                }
                if (tri_next < tri_end) return true;
                if (ANY_HIT && h.found && h.t < safe_t) {
                    occluded = true;
                    return false;
                }
                cur_t = h.t;
                do_pop = true;
            }
        }

        if (!advance_instance && !do_pop) {
            // interior node of the current level: test both children
            const float4 *pa = nodes + 2 * first;
            float4 a_lo = pa[0], a_hi = pa[1], b_lo = pa[2], b_hi = pa[3];
            uint32_t a = first, b = first + 1u;
            float da = slab_pre(cur, cur_t, a_lo, a_hi);
            float db = slab_pre(cur, cur_t, b_lo, b_hi);
            if (da > db) {
                float tf = da; da = db; db = tf;
                uint32_t ti = a; a = b; b = ti;
                float4 t0 = a_lo; a_lo = b_lo; b_lo = t0;
                float4 t1 = a_hi; a_hi = b_hi; b_hi = t1;
            }
            if (da == kTFar) {
                do_pop = true;
            } else {
                node_first = __float_as_uint(a_lo.w);
                node_count = __float_as_uint(a_hi.w);
                if (db != kTFar) {
                    if (sp < stk.capacity) {
                        stk.put(sp, pack_entry(b, __float_as_uint(b_lo.w), __float_as_uint(b_hi.w)));
                        sp += 1;
                    } else {
                        atomicOr(sc.error_flags, kErrStackOverflow);
                    }
                }
                return true;
            }
        }

        if (do_pop) {
            if (sp > sp_base) {
                sp -= 1;
                pop_node(stk.get(sp));
                return true;
            }
            if (!in_blas) return false;  // TLAS stack empty: done
            advance_instance = true;
        }
        want_advance = true;
        return true;
    }

    // Finishes the BLAS in progress (if any), then enters the next instance
    // of the TLAS leaf or returns to the TLAS.
    __device__ __forceinline__ bool advance(const DevScene &sc, LdsStack &stk) {
        if (in_blas) {
            const pulse_mesh_instance &mi = sc.instances[inst];
            if (!h.found) h.t = kTFar;
            float ow[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) ow[i] = mi.object_world[i];
            V3 hit_world = transform_position(ow, add(cur.o, scale(h.t, cur.d)));
            float new_t = length(sub(hit_world, wo));
            if (ANY_HIT) {
                if (new_t < threshold) {
                    occluded = true;
                    return false;
                }
            } else if (new_t < rec_t) {
                // a miss whose world t is finite records the reference's
                // unhit object record: instance 0, triangle 0
                // (utilities.wgsl:265)
                rec_t = new_t;
                rec_inst = h.found ? inst : 0u;
                rec_tri = h.tri;
                rec_u = h.u;
                rec_v = h.v;
                world_cull = rec_t;
            }
        }
        if (leaf_next < leaf_end) {
            inst = sc.instance_indices[leaf_next];
            leaf_next += 1;
            const pulse_mesh_instance &mi = sc.instances[inst];
            float wm[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) wm[i] = mi.world_object[i];
            uint32_t node_off = mi.node_offset;
            idx_off = mi.index_offset;
            V3 oo = transform_position(wm, wo);
            V3 dd = transform_direction(wm, wd);
            dlen = length(dd);
            cur = prep_ray(oo, divs(dd, dlen), sc.fast_div);
            // conservative object-space cull bound (see world_bound_slack)
            start_t = kTFar;
            if (world_cull < kTFar) {
                float wb = world_cull + world_bound_slack(wo, world_cull);
                float ob = wb * dlen * (1.0f + 0x1p-10f);
                if (ob < kTFar) start_t = ob;
            }
            if (ANY_HIT) {
                float sw = threshold - slack_thr;
                safe_t = sw > 0.0f ? sw * dlen * (1.0f - 0x1p-10f) : -1.0f;
            }
            h.t = start_t;
            h.found = false;
            h.tri = 0;
            h.u = 0.0f;
            h.v = 0.0f;
            cur_t = h.t;
            nodes = sc.blas_nodes + 2 * node_off;
            load_node(nodes, 0u);
            in_blas = true;
            sp_base = sp;
            iter_b = 0;
            tri_next = 0;
            tri_end = 0;
            return true;
        }
        // back to the TLAS
        in_blas = false;
        cur = world;
        cur_t = world_cull;
        nodes = sc.tlas_nodes;
        sp_base = 0;
        if (sp == 0) return false;
        sp -= 1;
        pop_node(stk.get(sp));
        return true;
    }

    // Stack entries: a pushed node carries its (first, count) words when
    // they fit, (count + 1) << 27 | first (first < 2^27, count < 31), so a
    // pop needs no memory access; otherwise its index (top 5 bits zero: the
    // host keeps node arrays below 2^27 entries) and the pop loads them.
    __device__ __forceinline__ static uint32_t pack_entry(uint32_t index, uint32_t first, uint32_t count) {
        return (first < (1u << 27) && count < 31u) ? ((count + 1u) << 27) | first : index;
    }
    __device__ __forceinline__ void pop_node(uint32_t e) {
        if (e >> 27) {
            node_first = e & ((1u << 27) - 1u);
            node_count = (e >> 27) - 1u;
        } else {
            load_node(nodes, e);
        }
    }

    // The child-pair index / first primitive and the primitive count of node
    // i (words 3 and 7 of the 32-byte node; its box was tested by the parent).
    __device__ __forceinline__ void load_node(const float4 *base, uint32_t i) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(base + 2 * i);
        node_first = w[3];
        node_count = w[7];
    }
};

// Closest hit (ANY_HIT = false): rec_* updated in place, returns false.
// Occlusion (ANY_HIT = true): returns true iff the reference's closest-hit t
// would be < threshold (see Traversal).
template <bool ANY_HIT>
__device__ __forceinline__ bool trace_unified(const DevScene &sc, V3 wo, V3 wd, float &rec_t, uint32_t &rec_inst,
                                              uint32_t &rec_tri, float &rec_u, float &rec_v, float threshold,
                                              LdsStack &stk) {
    Traversal<ANY_HIT> tr;
    if (tr.begin(sc, wo, wd, rec_t, threshold)) {
        while (tr.step(sc, stk)) {
        }
    }
    if (ANY_HIT) return tr.occluded;
    rec_t = tr.rec_t;
    rec_inst = tr.rec_inst;
    rec_tri = tr.rec_tri;
    rec_u = tr.rec_u;
    rec_v = tr.rec_v;
    return false;
}

}  // namespace pt
