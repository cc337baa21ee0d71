// wide.h -- two-level traversal over the 4-wide BVHs derived at upload
// (host/wide_bvh.h has the node layout).
//
// Same interface and the same closest-hit / occlusion definitions as
// Traversal (traverse.h), fewer and wider steps: one step fetches one
// 64-byte quantised node, tests its four child boxes, visits the nearest hit
// child next and pushes the others (farthest first).  The triangle test, the
// per-instance object-space closest hit and its world-t conversion are the
// reference's (utilities.wgsl:259-326).  The reference's BLAS leaves of more
// than two triangles are split at upload into leaves of at most two with
// tighter boxes (host/wide_bvh.h: refine_leaves), and of two hits at exactly
// equal t in one reference leaf the lower slot is kept, as the reference's
// slot-order walk of that leaf keeps it; what differs from the reference is
// the order in which its leaves are reached.  That order decides which of two
// hits with exactly equal t in different reference leaves is kept (the first
// found), and a box whose computed entry distance is rounded above a hit
// inside it; the quantised boxes contain the exact ones and the box test is
// padded by more than its rounding error, so no box holding a hit below the
// current bound is skipped.
#pragma once

#include "pt_device.h"
#include "traverse.h"

namespace pt {

constexpr uint32_t kWideEmpty = 0xffffffffu;
constexpr uint32_t kRangeMask = (1u << 26) - 1u;
constexpr uint32_t kRangeOne = 1u << 26;            // one primitive remaining
constexpr uint32_t kRangeStep = 1u - (1u << 26);    // next + 1, remaining - 1

// Origin, direction and clamped reciprocal direction (components below 2^-80
// in magnitude count as 2^-80: the slab of a parallel axis stays finite).
struct WideRay {
    V3 o, d, r;
};

// The hardware reciprocal (within 1 ulp): its error is covered by the
// slab test's padding.
__device__ __forceinline__ float wide_rcp(float v) {
    return __builtin_amdgcn_rcpf(fabsf(v) < 0x1p-80f ? copysignf(0x1p-80f, v) : v);
}
__device__ __forceinline__ WideRay wide_ray(V3 o, V3 d) {
    WideRay w;
    w.o = o;
    w.d = d;
    w.r = mk(wide_rcp(d.x), wide_rcp(d.y), wide_rcp(d.z));
    return w;
}

// Ray-side constants of one quantised node: along each axis the distance to
// plane q is q * (scale / d) + (origin - o) / d, one fma per plane.  The
// plane bytes enter that fma as the binary16 values 1024 + q (qpair below),
// so the constant term is (origin - o) / d - 1024 scale / d.  The rounding of
// those distances (the reciprocal, the products, the constant term, the fma)
// is a few 2^-24 of |origin - o| / |d| + 2048 scale / |d| on that axis; the
// constant term of the entry plane is lowered and that of the exit plane
// raised by 2^-20 of it, so every computed slab contains the exact one.
// Which plane of an axis is the entry is fixed by the sign of d on that
// axis (the low plane when d >= 0), so the entry and exit plane bytes are
// picked once per node instead of ordering each child's pair of distances:
// both give the same distances (fma rounding is monotone and the padded
// entry term never exceeds the exit term).
struct QuantFrame {
    float srx, sry, srz;                 // scale / d per axis
    float bnx, bfx, bny, bfy, bnz, bfz;  // (origin - o) / d - 1024 scale / d for the entry / exit planes, padded
};
__device__ __forceinline__ void quant_axis(float s, float org, float o, float r, float &sr, float &bn, float &bf) {
    sr = s * r;
    const float b = (org - o) * r;
    const float b1 = __builtin_fmaf(-1024.0f, sr, b);
    const float a = __builtin_fmaf(2048.0f, fabsf(sr), fabsf(b));
    bn = __builtin_fmaf(-0x1p-20f, a, b1);  // b1 - 2^-20 a, one rounding
    bf = __builtin_fmaf(0x1p-20f, a, b1);
}
__device__ __forceinline__ QuantFrame quant_frame(const WideRay &w, uint4 a) {
    QuantFrame f;
    quant_axis(__uint_as_float((a.w & 0xffu) << 23), __uint_as_float(a.x), w.o.x, w.r.x, f.srx, f.bnx, f.bfx);
    quant_axis(__uint_as_float(((a.w >> 8) & 0xffu) << 23), __uint_as_float(a.y), w.o.y, w.r.y, f.sry, f.bny, f.bfy);
    quant_axis(__uint_as_float(((a.w >> 16) & 0xffu) << 23), __uint_as_float(a.z), w.o.z, w.r.z, f.srz, f.bnz, f.bfz);
    return f;
}

// Two plane bytes of a plane word (children 0-1, or 2-3 with `hi`) as the
// two binary16 halves 1024 + q of one word: byte q under the exponent byte
// 0x64 is exactly 1024 + q (one byte permute, no conversions; the fma that
// uses a half widens it to f32 itself, v_fma_mix_f32).
__device__ __forceinline__ uint32_t qpair(uint32_t w, bool hi) {
    return __builtin_amdgcn_perm(0x64646464u, w, hi ? 0x04030402u : 0x04010400u);
}

// Planes of the four children as binary16 pairs: entry (n*) and exit (f*)
// plane per axis, children 0-1 in [0] and 2-3 in [1].
struct QuantPlanes {
    uint32_t nx[2], fx[2], ny[2], fy[2], nz[2], fz[2];
};
__device__ __forceinline__ void quant_axis_planes(bool neg, uint32_t lo, uint32_t hi, uint32_t *n, uint32_t *f) {
    const uint32_t wn = neg ? hi : lo, wf = neg ? lo : hi;
    n[0] = qpair(wn, false);
    n[1] = qpair(wn, true);
    f[0] = qpair(wf, false);
    f[1] = qpair(wf, true);
}
__device__ __forceinline__ QuantPlanes quant_planes(const WideRay &w, uint4 b, uint4 c) {
    QuantPlanes p;
    quant_axis_planes(w.r.x < 0.0f, b.x, b.y, p.nx, p.fx);
    quant_axis_planes(w.r.y < 0.0f, b.z, b.w, p.ny, p.fy);
    quant_axis_planes(w.r.z < 0.0f, c.x, c.y, p.nz, p.fz);
    return p;
}

// 1024 + the plane byte of child k.
__device__ __forceinline__ float qplane(const uint32_t *pair, int k) {
    const uint32_t w = pair[k >> 1];
    const uint16_t bits = static_cast<uint16_t>((k & 1) ? w >> 16 : w & 0xffffu);
    return static_cast<float>(__builtin_bit_cast(_Float16, bits));
}

// Slab test of child k: its entry distance, or kTFar when the box is
// missed, behind the origin or at or beyond `bound`.
__device__ __forceinline__ float quant_slab(const QuantFrame &f, float bound, const QuantPlanes &p, int k) {
    const float tmin = fmaxf(fmaxf(__builtin_fmaf(qplane(p.nx, k), f.srx, f.bnx), __builtin_fmaf(qplane(p.ny, k), f.sry, f.bny)),
                             __builtin_fmaf(qplane(p.nz, k), f.srz, f.bnz));
    const float tmax = fminf(fminf(__builtin_fmaf(qplane(p.fx, k), f.srx, f.bfx), __builtin_fmaf(qplane(p.fy, k), f.sry, f.bfy)),
                             __builtin_fmaf(qplane(p.fz, k), f.srz, f.bfz));
    return (tmax >= tmin && tmin < bound && tmax > 0.0f) ? tmin : kTFar;
}

// Moeller-Trumbore exactly as mt() (pt_device.h) evaluates it, without its
// early exits: every quantity is computed and the hit is taken when all of
// mt()'s tests pass, so lanes of a wave stay together through the test.
// `tri` rows are wide_tri's: the third float4 holds e2.z, the triangle's
// tri_bvh slot and the end slot of the reference's leaf holding it.  h.tri
// receives the tri_bvh slot.  With TIE_IN_LEAF a hit at exactly the current
// t also replaces the current one when both lie in one reference leaf and
// its slot is the lower: the reference tests a leaf in slot order and keeps
// the first of equal hits, and a refined leaf's triangles come in another
// order.
template <bool TIE_IN_LEAF>
__device__ __forceinline__ void mt_select(V3 o, V3 d, const float4 *tri, uint32_t slot, ObjHit &h) {
    const float4 A = tri[3 * slot];
    const float4 B = tri[3 * slot + 1];
    const float4 C = tri[3 * slot + 2];
    const uint32_t ref = __float_as_uint(C.y), leaf_end = __float_as_uint(C.z);
    const V3 p0 = mk(A.x, A.y, A.z);
    const V3 e1 = mk(A.w, B.x, B.y);
    const V3 e2 = mk(B.z, B.w, C.x);
    const V3 hv = cross(d, e2);
    const float a = dot(e1, hv);
    // rcp_newton is 1.0f / a wherever it can decide a hit: the 4-wide BVHs
    // exist only for scenes whose edge components are finite and at most
    // 2^60 (so |a| < 2^124 for the unit direction), and |a| < 1e-4, inf and
    // NaN give no hit with either value
    const float f = rcp_newton(a);
    const V3 s = sub(o, p0);
    const float u = f * dot(s, hv);
    const V3 q = cross(s, e1);
    const float v = f * dot(d, q);
    const float t = f * dot(e2, q);
    const bool nearer = t < h.t || (TIE_IN_LEAF && t == h.t && ref < h.tri && h.tri < leaf_end);
    const bool hit = !(fabsf(a) < 0.0001f) && !(u < 0.0f || u > 1.0f) && !(v < 0.0f || u + v > 1.0f) &&
                     t > 0.001f && nearer;
    if (hit) {
        h.t = t;
        h.tri = ref;
        h.u = u;
        h.v = v;
        h.found = true;
    }
}

// transform_position / transform_direction (utilities.wgsl:751-759) with a
// matrix whose bottom row is 0 0 0 1, from its columns 0-3, rows 0-2 packed
// in three 16-byte words (an instance record line): mat_vec's sums in its
// order, with w = 1 (m[12 + r] * 1 is m[12 + r]) or w = 0.
__device__ __forceinline__ V3 affine_point(uint4 a, uint4 b, uint4 c, V3 p) {
    const float m0 = __uint_as_float(a.x), m1 = __uint_as_float(a.y), m2 = __uint_as_float(a.z);
    const float m4 = __uint_as_float(a.w), m5 = __uint_as_float(b.x), m6 = __uint_as_float(b.y);
    const float m8 = __uint_as_float(b.z), m9 = __uint_as_float(b.w), m10 = __uint_as_float(c.x);
    const float m12 = __uint_as_float(c.y), m13 = __uint_as_float(c.z), m14 = __uint_as_float(c.w);
    return mk(m0 * p.x + m4 * p.y + m8 * p.z + m12, m1 * p.x + m5 * p.y + m9 * p.z + m13,
              m2 * p.x + m6 * p.y + m10 * p.z + m14);
}
__device__ __forceinline__ V3 affine_direction(uint4 a, uint4 b, uint4 c, V3 p) {
    const float m0 = __uint_as_float(a.x), m1 = __uint_as_float(a.y), m2 = __uint_as_float(a.z);
    const float m4 = __uint_as_float(a.w), m5 = __uint_as_float(b.x), m6 = __uint_as_float(b.y);
    const float m8 = __uint_as_float(b.z), m9 = __uint_as_float(b.w), m10 = __uint_as_float(c.x);
    const float m12 = __uint_as_float(c.y), m13 = __uint_as_float(c.z), m14 = __uint_as_float(c.w);
    return mk(m0 * p.x + m4 * p.y + m8 * p.z + m12 * 0.0f, m1 * p.x + m5 * p.y + m9 * p.z + m13 * 0.0f,
              m2 * p.x + m6 * p.y + m10 * p.z + m14 * 0.0f);
}

__device__ __forceinline__ void wide_cswap(float &da, uint32_t &ea, float &db, uint32_t &eb) {
    const bool s = db < da;
    const float d = s ? db : da;
    const uint32_t e = s ? eb : ea;
    db = s ? da : db;
    eb = s ? ea : eb;
    da = d;
    ea = e;
}

// No iteration guard: every interior entry of a 4-wide tree names a node of
// a larger index (the collapse allocates a node before its children, and the
// upload refuses a tree where that fails: wide_bvh.h, collapse_wide), so a
// walk visits each node of a tree at most once per entry into the tree and
// ends.  COUNT keeps the work census (node visits, triangle tests,
// instances entered).
template <bool ANY_HIT, int TRIS_PER_STEP = 2, bool COUNT = false>
struct WideTraversal {
    V3 wo, wd;         // world ray
    float threshold, slack_thr, any_cull;  // any_cull: occlusion mode's world culling bound
    WideRay cur;       // ray of the current level
    const uint4 *level;  // 16-byte words of the current level's tree (its root node first)
    bool in_blas;
    int sp, sp_base;
    int iter;             // wide node visits of this ray (work census)
    // primitive ranges in progress as next | remaining << 26 (one add steps
    // both): instances of a TLAS leaf, triangles of a BLAS leaf
    uint32_t leaf, tri;
    uint32_t inst;               // instance record (instance_indices position) in progress
    bool want_advance;           // next: leave the BLAS and/or enter the next instance
    float safe_t;
    ObjHit h;
    uint32_t entry;              // current entry (wide node or leaf range)
    uint32_t n_tri, n_inst;      // triangle tests, instances entered (work census; unread otherwise)
    float rec_t, rec_u, rec_v;
    uint32_t rec_inst, rec_tri;
    bool occluded;

    // World-space culling bound: the closest hit so far, or the padded
    // occlusion threshold.
    __device__ __forceinline__ float wcull() const { return ANY_HIT ? any_cull : rec_t; }

    // false: nothing to traverse (empty TLAS); the result is already final.
    __device__ __forceinline__ bool begin(const DevScene &sc, V3 o, V3 d, float t_max, float thr) {
        rec_t = t_max;
        rec_inst = 0;
        rec_tri = 0;
        rec_u = 0.0f;
        rec_v = 0.0f;
        occluded = false;
        const uint32_t *root = reinterpret_cast<const uint32_t *>(sc.tlas_nodes);
        if (root[3] == 0u && root[7] == 0u) return false;  // utilities.wgsl:129-131
        wo = o;
        wd = d;
        threshold = thr;
        slack_thr = 0.0f;
        any_cull = 0.0f;
        if (ANY_HIT) {
            slack_thr = world_bound_slack(wo, threshold);
            any_cull = threshold + slack_thr;
        }
        cur = wide_ray(wo, wd);
        level = reinterpret_cast<const uint4 *>(sc.wide_tlas);
        entry = 0u;
        in_blas = false;
        sp = 0;
        sp_base = 0;
        iter = 0;
        n_tri = 0;
        n_inst = 0;
        leaf = 0;
        tri = 0;
        want_advance = false;
        inst = 0;
        safe_t = -1.0f;
        h.t = kTFar;
        h.found = false;
        h.tri = 0;
        h.u = 0.0f;
        h.v = 0.0f;
        return true;
    }

    // One step; false when the traversal is complete (run_advance as in
    // Traversal::step).  A step tests the current interior node (if any) and
    // then, when that leads to a leaf or the current entry is one, up to
    // TRIS_PER_STEP of the leaf's triangles: a lane that reaches a leaf does
    // not wait a step to start on it, so more lanes of a wave work in each
    // of the two parts.
    __device__ __forceinline__ bool step(const DevScene &sc, LdsStack &stk, bool run_advance = true) {
        if (want_advance) {
            if (!run_advance) return true;
            want_advance = false;
            return advance(sc, stk);
        }
        if (!(in_blas && tri >= kRangeOne) && !(entry >> 31)) {
            // interior wide node of the current level
            if (COUNT) iter += 1;
            const float cur_t = in_blas ? h.t : wcull();  // culling bound of this level
            const uint4 *p = level + 4u * entry;
            const uint4 qa = p[0], qb = p[1], qc = p[2], qd = p[3];
            const QuantFrame f = quant_frame(cur, qa);
            const QuantPlanes pl = quant_planes(cur, qb, qc);
            uint32_t e0 = qc.z, e1 = qc.w, e2 = qd.x, e3 = qd.y;
            // all four boxes are tested (an empty position's planes are
            // zero), then empty positions are dropped
            float d0 = quant_slab(f, cur_t, pl, 0), d1 = quant_slab(f, cur_t, pl, 1);
            float d2 = quant_slab(f, cur_t, pl, 2), d3 = quant_slab(f, cur_t, pl, 3);
            d0 = e0 != kWideEmpty ? d0 : kTFar;
            d1 = e1 != kWideEmpty ? d1 : kTFar;
            d2 = e2 != kWideEmpty ? d2 : kTFar;
            d3 = e3 != kWideEmpty ? d3 : kTFar;
            wide_cswap(d0, e0, d1, e1);
            wide_cswap(d2, e2, d3, e3);
            wide_cswap(d0, e0, d2, e2);
            wide_cswap(d1, e1, d3, e3);
            wide_cswap(d1, e1, d2, e2);
            if (d0 == kTFar) {
                if (!pop(sc, stk)) return false;
                if (want_advance) return true;
            } else {
                entry = e0;
                push3(sc, stk, d1, e1, d2, e2, d3, e3);
            }
        }
        if (!(in_blas && tri >= kRangeOne)) {
            const uint32_t e = entry;
            if (!(e >> 31)) return true;  // an interior node: the next step
            const uint32_t first = e & ((1u << 26) - 1u);
            const uint32_t count = ((e >> 26) & 31u) + 1u;
            if (!in_blas) {
                leaf = first | (count << 26);
                inst = 0xffffffffu;
                want_advance = true;
                return true;
            }
            tri = first | (count << 26);
        }
        // BLAS leaf entries hold global slots of wide_tri; h.tri (a global
        // tri_bvh slot) becomes the mesh's slot at the merge
        mt_select<!ANY_HIT>(cur.o, cur.d, sc.wide_tri, tri & kRangeMask, h);
        tri += kRangeStep;
        if (COUNT) n_tri += 1;
        if (TRIS_PER_STEP == 2 && tri >= kRangeOne) {
            mt_select<!ANY_HIT>(cur.o, cur.d, sc.wide_tri, tri & kRangeMask, h);
            tri += kRangeStep;
            if (COUNT) n_tri += 1;
        }
        if (tri >= kRangeOne) return true;
        if (ANY_HIT && h.found && h.t < safe_t) {
            occluded = true;
            return false;
        }
        return pop(sc, stk);
    }

    // Pushes the hit ones of the sorted children 1-3, farthest first.  They
    // are a prefix of that order (a miss sorts last), so with three free
    // LDS entries the pushes are three unconditional writes (the unused ones
    // above the new top are never read) and one add; else one push each.
    __device__ __forceinline__ void push3(const DevScene &sc, LdsStack &stk, float d1, uint32_t e1, float d2,
                                          uint32_t e2, float d3, uint32_t e3) {
        if (sp + 3 <= stk.lds_cap && sp + 3 <= stk.capacity) {
            const int n = (d1 != kTFar ? 1 : 0) + (d2 != kTFar ? 1 : 0) + (d3 != kTFar ? 1 : 0);
            stk.lds()[sp * stk.stride] = n == 3 ? e3 : (n == 2 ? e2 : e1);
            stk.lds()[(sp + 1) * stk.stride] = n == 3 ? e2 : e1;
            stk.lds()[(sp + 2) * stk.stride] = e1;
            sp += n;
        } else {
            push(sc, stk, d3, e3);
            push(sc, stk, d2, e2);
            push(sc, stk, d1, e1);
        }
    }

    __device__ __forceinline__ void push(const DevScene &sc, LdsStack &stk, float d, uint32_t e) {
        if (d == kTFar) return;
        if (sp < stk.capacity) {
            stk.put(sp, e);
            sp += 1;
        } else {
            atomicOr(sc.error_flags, kErrStackOverflow);
        }
    }

    __device__ __forceinline__ bool pop(const DevScene &sc, LdsStack &stk) {
        if (sp > sp_base) {
            sp -= 1;
            entry = stk.get(sp);
            return true;
        }
        if (!in_blas) return false;  // TLAS stack empty: done
        want_advance = true;
        return true;
    }

    // Finishes the BLAS in progress (if any), then enters the next instance
    // of the TLAS leaf or returns to the TLAS.
    __device__ __forceinline__ bool advance(const DevScene &sc, LdsStack &stk) {
        // A BLAS without a hit below the bound skips the merge when its
        // world t through the miss point (t = 1e30) certainly overflows to
        // +inf and so never updates (utilities.wgsl:318-325): bit 31 of inst
        // is set on entry when that is not certain (miss_overflows).
        if (in_blas && (h.found || (inst >> 31))) {
            // the record's second line: object_world, index_offset, instance
            const uint4 *rb = sc.wide_inst + 8u * (inst & kRangeMask) + 4u;
            const uint4 b0 = rb[0], b1 = rb[1], b2 = rb[2], b3 = rb[3];
            if (!h.found) h.t = kTFar;
            const V3 hit_world = affine_point(b0, b1, b2, add(cur.o, scale(h.t, cur.d)));
            float new_t = length(sub(hit_world, wo));
            if (ANY_HIT) {
                if (new_t < threshold) {
                    occluded = true;
                    return false;
                }
            } else if (new_t < rec_t) {
                // a miss whose world t is finite (an instance that scales
                // down by ~1e-11 or more) records the reference's unhit
                // object record: instance 0, triangle 0 (utilities.wgsl:265)
                rec_t = new_t;
                rec_inst = h.found ? b3.y : 0u;
                rec_tri = h.found ? h.tri - b3.x : 0u;
                rec_u = h.u;
                rec_v = h.v;
            }
        }
        if (leaf >= kRangeOne) {
            inst = leaf & kRangeMask;
            leaf += kRangeStep;
            if (COUNT) n_inst += 1;
            // the record's first line: world_object and the BLAS root
            const uint4 *ra = sc.wide_inst + 8u * inst;
            const uint4 a0 = ra[0], a1 = ra[1], a2 = ra[2], a3 = ra[3];
            if (!(a3.w != 0u && fmaxf(fmaxf(fabsf(wo.x), fabsf(wo.y)), fabsf(wo.z)) < 0x1p60f)) inst |= 0x80000000u;
            V3 oo = affine_point(a0, a1, a2, wo);
            V3 dd = affine_direction(a0, a1, a2, wd);
            const float dlen = length(dd);
            cur = wide_ray(oo, divs(dd, dlen));
            float start_t = kTFar;
            if (wcull() < kTFar) {
                float wb = wcull() + world_bound_slack(wo, wcull());
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
            level = reinterpret_cast<const uint4 *>(sc.wide_blas) + 4u * a3.x;
            in_blas = true;
            sp_base = sp;
            // (no triangles remain here: every leaf runs to its end)
            entry = 0u;
            return true;
        }
        // back to the TLAS
        in_blas = false;
        cur = wide_ray(wo, wd);
        level = reinterpret_cast<const uint4 *>(sc.wide_tlas);
        sp_base = 0;
        if (sp == 0) return false;
        sp -= 1;
        entry = stk.get(sp);
        return true;
    }
};

// Closest hit / occlusion of one ray over the 4-wide BVHs (trace_unified's
// contract).
template <bool ANY_HIT>
__device__ __forceinline__ bool trace_wide(const DevScene &sc, V3 wo, V3 wd, float &rec_t, uint32_t &rec_inst,
                                           uint32_t &rec_tri, float &rec_u, float &rec_v, float threshold,
                                           LdsStack &stk) {
    WideTraversal<ANY_HIT> tr;
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
