// pt_device.h -- device-side building blocks of the gfx950 path-tracing hot
// path: scene view, f32 vector helpers, PCG, slab and Moeller-Trumbore tests,
// two-level (TLAS -> BLAS) stack traversal with an LDS stack, GGX sampling
// and next-event-estimation terms.
//
// Every arithmetic expression follows the reference WGSL operation by
// operation (src/utilities.wgsl, src/path_tracer/path_tracer.wgsl) and the
// library is compiled with -ffp-contract=off, so results are those of the
// reference's f32 arithmetic with IEEE division/sqrt; transcendentals come
// from pt_math.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pulse_hip.h"

#include "dual.h"

namespace pt {

constexpr float kPI = 3.14159265358f;     // utilities.wgsl:32
constexpr float kTWO_PI = 6.28318530718f; // utilities.wgsl:34
constexpr float kINV_PI = 0.31830988618f; // utilities.wgsl:35
constexpr float kTFar = 1e30f;            // path_tracer.wgsl:74
constexpr int kMaxIterations = 100000;    // utilities.wgsl:137, 271

constexpr uint32_t kErrStackOverflow = 1u;
constexpr uint32_t kErrIterationCap = 2u;

// Device view of the uploaded scene.  Besides the reference buffers it holds
// two traversal-ordered copies made at upload time (identical values, no
// index indirection in the inner loop):
//   tri_bvh[3*k .. 3*k+2]  = {p0.xyz, e1.x}, {e1.y, e1.z, e2.x, e2.y}, {e2.z,0,0,0}
//   nrm_bvh[3*k .. 3*k+2]  = n_first, n_second, n_third
// for triangle-index slot k, with e1 = p1 - p0 and e2 = p2 - p0 evaluated in
// f32 exactly as ray_triangle_intersect does (utilities.wgsl:432-433).
struct DevScene {
    const float4 *tri_bvh;
    const float4 *nrm_bvh;
    const pulse_primitive *primitives;  // original order (NEE, utilities.wgsl:511)
    const float4 *blas_nodes;           // 2 float4 per node
    const float4 *tlas_nodes;           // 2 float4 per node
    const uint32_t *instance_indices;
    const pulse_mesh_instance *instances;
    const pulse_material *materials;
    const float *light_cdfs;
    const float *light_areas;
    const pulse_light_data_index *lights;
    const float *strength_cdf;          // light_emission_strength_cdf (may be short: checked on the host)
    const float4 *wide_tlas;            // 4-wide TLAS, 4 float4 per quantised node (null: not built)
    const float4 *wide_blas;            // all meshes' 4-wide BLASes
    const uint4 *wide_inst;             // per instance_indices entry: 128-byte instance record (below)
    const float4 *wide_tri;             // tri_bvh rows, refined BLAS leaves reordered (wide.h: mt_select)
    uint32_t instance_count;
    uint32_t light_count;
    uint32_t *error_flags;
    // every node coordinate is 0 or has magnitude in [2^-56, 2^40] (the
    // range in which the slab test may divide by reciprocal multiplication)
    bool fast_div;
};

// Instance record of the 4-wide traversal, one per instance_indices entry
// (the 4-wide TLAS leaves index these records directly), 32 words in two
// 64-byte lines: the first is read when a ray enters the instance, the
// second when it leaves it.  Columns 0-3, rows 0-2 of the column-major
// matrices: the bottom row of every matrix is 0 0 0 1 (the 4-wide BVHs are
// built only then), so transform_position's w is 1 and the 3x4 part gives
// its result bit for bit.
constexpr int kInstRecordWords = 32;
constexpr int kRecWorldObject = 0;   // words 0-11: world_object
constexpr int kRecRoot = 12;         // first wide node of the instance's BLAS (words 13-14 zero)
constexpr int kRecMissOverflows = 15;  // 1: a miss's world t overflows for rays from |o| < 2^60 (miss_overflows)
constexpr int kRecObjectWorld = 16;  // words 16-27: object_world
constexpr int kRecIndexOffset = 28;  // the instance's index_offset
constexpr int kRecInstance = 29;     // the instance (instances[] index); words 30-31 zero
// The matrix words of a record: columns 0-3, rows 0-2 of m (16 floats,
// column-major) at rec[0 .. 11].
__host__ __device__ inline void record_matrix(const float *m, uint32_t *rec) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) {
            float v = m[4 * c + r];
            uint32_t u;
            __builtin_memcpy(&u, &v, 4);
            rec[3 * c + r] = u;
        }
}
// True when, for an affine instance with matrices m (object_world) and w
// (world_object), the reference's world t of a BLAS miss -- the length of
// m (o' + 1e30 d') - o with o' = w o, d' the unit object direction --
// overflows to +inf for every ray origin o with components below 2^60, so a
// miss can never update the closest hit (utilities.wgsl:318-325).  With the
// 3x3 parts' Frobenius norms ||w|| <= 2^30 and ||m|| ||w|| <= 2^16 and a
// translation below 2^40: |o'| < 2^92, so the miss point's difference from
// o is 1e30 m d' up to rounding, of length >= 1e30 / ||w|| >= 2^69, whose
// largest component squared alone overflows; the rounding of the f32
// products stays below 2^-6 of it.
__host__ __device__ inline bool miss_overflows(const float *m, const float *w) {
    double fm = 0.0, fw = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            fm += static_cast<double>(m[4 * c + r]) * m[4 * c + r];
            fw += static_cast<double>(w[4 * c + r]) * w[4 * c + r];
        }
    const bool t_ok = fabs(m[12]) <= 0x1p40f && fabs(m[13]) <= 0x1p40f && fabs(m[14]) <= 0x1p40f;
    return t_ok && fw <= 0x1p60 && fm * fw <= 0x1p32;
}
// A bottom row of 0 0 0 1: transform_position's w is then 1 for every
// finite point.
__host__ __device__ inline bool affine_matrix(const float *m) {
    return m[3] == 0.0f && m[7] == 0.0f && m[11] == 0.0f && m[15] == 1.0f;
}

// ---------------------------------------------------------------------------
// 3-vectors with the reference's evaluation order.  Templated on the scalar:
// float in the forward kernels, Dual<N> in the adjoint replay (dual.h), whose
// values follow exactly the same f32 operations.
// ---------------------------------------------------------------------------
template <class S>
struct V3T {
    S x, y, z;
};
using V3 = V3T<float>;
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
template <class S>
__device__ __forceinline__ V3T<S> mk3(S x, S y, S z) { return V3T<S>{x, y, z}; }
template <class A, class B>
__device__ __forceinline__ auto add(const V3T<A> &a, const V3T<B> &b) -> V3T<decltype(a.x + b.x)> {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class A, class B>
__device__ __forceinline__ auto sub(const V3T<A> &a, const V3T<B> &b) -> V3T<decltype(a.x - b.x)> {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class S>
__device__ __forceinline__ V3T<S> neg(const V3T<S> &a) { return {-a.x, -a.y, -a.z}; }
template <class A, class B>
__device__ __forceinline__ auto mul(const V3T<A> &a, const V3T<B> &b) -> V3T<decltype(a.x * b.x)> {
    return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class A, class B>
__device__ __forceinline__ auto scale(const A &s, const V3T<B> &a) -> V3T<decltype(s * a.x)> {
    return {s * a.x, s * a.y, s * a.z};
}
template <class A, class B>
__device__ __forceinline__ auto divs(const V3T<A> &a, const B &s) -> V3T<decltype(a.x / s)> {
    return {a.x / s, a.y / s, a.z / s};
}
template <class A, class B>
__device__ __forceinline__ auto dot(const V3T<A> &a, const V3T<B> &b) -> decltype(a.x * b.x) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class A, class B>
__device__ __forceinline__ auto cross(const V3T<A> &a, const V3T<B> &b) -> V3T<decltype(a.x * b.x)> {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class S>
__device__ __forceinline__ S length(const V3T<S> &a) { return sqrt_(dot(a, a)); }
template <class S>
__device__ __forceinline__ V3T<S> normalize(const V3T<S> &a) { return divs(a, length(a)); }
template <class S>
__device__ __forceinline__ S max3(const V3T<S> &v) { return fmax_(fmax_(v.x, v.y), v.z); }
__device__ __forceinline__ V3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
template <class S>
__device__ __forceinline__ V3T<S> lift(const V3 &a) { return {sconst<S>(a.x), sconst<S>(a.y), sconst<S>(a.z)}; }

// Column-major mat4 * (p, w) with ((c0*x + c1*y) + c2*z) + c3*w per row.
template <class S>
__device__ __forceinline__ void mat_vec(const float *m, const V3T<S> &p, float w, S h[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) h[r] = m[r] * p.x + m[4 + r] * p.y + m[8 + r] * p.z + m[12 + r] * w;
}
// transform_position / transform_direction, utilities.wgsl:751-759
template <class S>
__device__ __forceinline__ V3T<S> transform_position(const float *m, const V3T<S> &p) {
    S h[4];
    mat_vec(m, p, 1.0f, h);
    return {h[0] / h[3], h[1] / h[3], h[2] / h[3]};
}
// The same for f32, skipping the divide when w == 1 (x / 1 == x exactly),
// which is every affine transform applied to a finite point.
__device__ __forceinline__ V3 transform_position(const float *m, const V3 &p) {
    float h[4];
    mat_vec(m, p, 1.0f, h);
    if (h[3] == 1.0f) return {h[0], h[1], h[2]};
    return {h[0] / h[3], h[1] / h[3], h[2] / h[3]};
}
template <class S>
__device__ __forceinline__ V3T<S> transform_direction(const float *m, const V3T<S> &p) {
    S h[4];
    mat_vec(m, p, 0.0f, h);
    return {h[0], h[1], h[2]};
}

// ---------------------------------------------------------------------------
// R1: PCG hash, utilities.wgsl:67-83
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rand_u(uint32_t &state) {
    state = state * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float rand_f(uint32_t &state) {
    return static_cast<float>(rand_u(state)) * __uint_as_float(0x2f800004u);
}
__device__ __forceinline__ uint32_t pixel_seed(uint32_t pixel_index, uint32_t acc) {
    return pixel_index * 1235243u + acc * 5817321u;
}

// ---------------------------------------------------------------------------
// R5: slab test, utilities.wgsl:405-426 (true division per axis)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float slab(V3 o, V3 d, float rec_t, float4 lo, float4 hi) {
    float tx1 = (lo.x - o.x) / d.x;
    float tx2 = (hi.x - o.x) / d.x;
    float tmin = fminf(tx1, tx2);
    float tmax = fmaxf(tx1, tx2);
    float ty1 = (lo.y - o.y) / d.y;
    float ty2 = (hi.y - o.y) / d.y;
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = (lo.z - o.z) / d.z;
    float tz2 = (hi.z - o.z) / d.z;
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    return (tmax >= tmin && tmin < rec_t && tmax > 0.0f) ? tmin : kTFar;
}

// Object-space hit record of one BLAS traversal.
struct ObjHit {
    float t;
    uint32_t tri;
    float u, v;
    bool found;
};

// The hardware reciprocal (within 1 ulp) refined by one fused Newton step:
// equal to the correctly rounded division 1.0f / a bit for bit for every
// |a| in [2^-100, 2^126) (checked exhaustively on gfx950: both signs, every
// mantissa of every exponent in that range, 3.8e9 values); it differs where
// 1/a is subnormal (|a| >= 2^126).
__device__ __forceinline__ float rcp_newton(float a) {
    const float y0 = __builtin_amdgcn_rcpf(a);
    return __builtin_fmaf(__builtin_fmaf(-a, y0, 1.0f), y0, y0);
}
// 1.0f / a bit for bit: rcp_newton, and the division itself for the other
// a (zero, tiny, huge, inf, NaN).
__device__ __forceinline__ float rcp_rn(float a) {
    float y = rcp_newton(a);
    if (__builtin_expect(!(fabsf(a) >= 0x1p-100f && fabsf(a) < 0x1p126f), 0)) y = 1.0f / a;
    return y;
}

// R6: Moeller-Trumbore, utilities.wgsl:430-460, on a traversal-ordered
// triangle (p0, e1, e2).
__device__ __forceinline__ void mt(V3 o, V3 d, const float4 *tri, uint32_t slot, uint32_t tri_local, ObjHit &h) {
    float4 A = tri[3 * slot];
    float4 B = tri[3 * slot + 1];
    float4 C = tri[3 * slot + 2];
    V3 p0 = mk(A.x, A.y, A.z);
    V3 e1 = mk(A.w, B.x, B.y);
    V3 e2 = mk(B.z, B.w, C.x);
    V3 hv = cross(d, e2);
    float a = dot(e1, hv);
    if (fabsf(a) < 0.0001f) return;
    float f = rcp_rn(a);
    V3 s = sub(o, p0);
    float u = f * dot(s, hv);
    if (u < 0.0f || u > 1.0f) return;
    V3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return;
    float t = f * dot(e2, q);
    if (t > 0.001f && t < h.t) {
        h.t = t;
        h.tri = tri_local;
        h.u = u;
        h.v = v;
        h.found = true;
    }
}

// Per-thread traversal stack: entry i < lds_cap at base[i * stride] in LDS
// (consecutive lanes in consecutive banks); deeper entries, when a kernel
// keeps only the top of the stack in LDS to fit more waves per CU, at
// spill[(i - lds_cap) * spill_stride] in global memory (one column per lane,
// consecutive lanes adjacent).
struct LdsStack {
    uint32_t *base;
    int stride;
    int capacity;              // total entries
    int lds_cap = 1 << 30;     // entries held in LDS
    uint32_t *spill = nullptr;
    uint32_t spill_stride = 0;
    // `base` points into LDS (every stack's top part lives there): accessed
    // through an LDS pointer, an entry is an LDS read or write, never a
    // generic-address one that picks LDS or the spill column per lane.
    using lds_word = __attribute__((address_space(3))) uint32_t;
    __device__ __forceinline__ lds_word *lds() const { return (lds_word *)base; }
    __device__ __forceinline__ void put(int i, uint32_t v) {
        if (i < lds_cap) lds()[i * stride] = v;
        else spill[static_cast<uint32_t>(i - lds_cap) * spill_stride] = v;
    }
    __device__ __forceinline__ uint32_t get(int i) const {
        if (i < lds_cap) return lds()[i * stride];
        return spill[static_cast<uint32_t>(i - lds_cap) * spill_stride];
    }
};

// Slack of the object-space culling bound derived from a world-space bound:
// covers the rounding of the world->object->world round trip of the hit
// point that the reference's world-t conversion performs.
__device__ __forceinline__ float world_bound_slack(V3 o, float t) {
    float m = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fmaxf(fabsf(o.z), t));
    return m * 0x1p-12f + 0x1p-20f;
}

// R4: traverse_blas (utilities.wgsl:259-326) for one instance.  `world_bound`
// is an upper bound on the world t that can still matter (the current
// closest hit, or the occlusion threshold); it is turned into a conservative
// object-space bound so far nodes are culled, which cannot change which
// triangle wins (anything culled converts to a world t >= world_bound).
// Returns the reference's world t of this instance's closest object-space hit
// (+inf if none) and the object-space record.
__device__ __forceinline__ float traverse_blas(const DevScene &sc, V3 o, V3 d, uint32_t inst, float world_bound,
                                               LdsStack &stk, int sp_base, ObjHit &h) {
    const pulse_mesh_instance &mi = sc.instances[inst];
    float wo[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) wo[i] = mi.world_object[i];
    const uint32_t node_off = mi.node_offset;
    const uint32_t idx_off = mi.index_offset;
    V3 oo = transform_position(wo, o);
    V3 dd_raw = transform_direction(wo, d);
    float dlen = length(dd_raw);
    V3 od = divs(dd_raw, dlen);

    h.t = kTFar;
    h.found = false;
    h.tri = 0;
    h.u = 0.0f;
    h.v = 0.0f;
    if (world_bound < kTFar) {
        float wb = world_bound + world_bound_slack(o, world_bound);
        float ob = wb * dlen * (1.0f + 0x1p-10f);
        if (ob < kTFar) h.t = ob;
    }

    int sp = sp_base;
    uint32_t node = 0;
    float4 n_lo = sc.blas_nodes[2 * (node_off + 0)];
    float4 n_hi = sc.blas_nodes[2 * (node_off + 0) + 1];
    int iteration = 0;
    while (iteration < kMaxIterations) {
        iteration += 1;
        uint32_t first = __float_as_uint(n_lo.w);
        uint32_t count = __float_as_uint(n_hi.w);
        if (count > 0u) {
            for (uint32_t i = 0; i < count; ++i) mt(oo, od, sc.tri_bvh, idx_off + first + i, first + i, h);
            if (sp == sp_base) break;
            sp -= 1;
            node = stk.get(sp);
            n_lo = sc.blas_nodes[2 * (node_off + node)];
            n_hi = sc.blas_nodes[2 * (node_off + node) + 1];
            continue;
        }
        uint32_t a = first, b = first + 1u;
        const float4 *pa = sc.blas_nodes + 2 * (node_off + a);
        float4 a_lo = pa[0], a_hi = pa[1], b_lo = pa[2], b_hi = pa[3];
        float da = slab(oo, od, h.t, a_lo, a_hi);
        float db = slab(oo, od, h.t, b_lo, b_hi);
        if (da > db) {
            float tf = da; da = db; db = tf;
            uint32_t ti = a; a = b; b = ti;
            float4 t0 = a_lo; a_lo = b_lo; b_lo = t0;
            float4 t1 = a_hi; a_hi = b_hi; b_hi = t1;
        }
        if (da == kTFar) {
            if (sp == sp_base) break;
            sp -= 1;
            node = stk.get(sp);
            n_lo = sc.blas_nodes[2 * (node_off + node)];
            n_hi = sc.blas_nodes[2 * (node_off + node) + 1];
        } else {
            node = a;
            n_lo = a_lo;
            n_hi = a_hi;
            if (db != kTFar) {
                if (sp < stk.capacity) {
                    stk.put(sp, b);
                    sp += 1;
                } else {
                    atomicOr(sc.error_flags, kErrStackOverflow);
                }
            }
        }
    }
    if (iteration >= kMaxIterations) atomicOr(sc.error_flags, kErrIterationCap);

    // World t through the hit point (utilities.wgsl:318-320).  With no hit
    // the reference converts t = 1e30, which overflows to +inf.
    if (!h.found) h.t = kTFar;
    float ow[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) ow[i] = mi.object_world[i];
    V3 hit_obj = add(oo, scale(h.t, od));
    V3 hit_world = transform_position(ow, hit_obj);
    return length(sub(hit_world, o));
}

// R3: traverse_tlas (utilities.wgsl:127-188).  Closest hit: rec_t/ids are
// updated in place.  Shadow (any_hit) mode: returns true as soon as an
// instance's world t is below `threshold` (the reference's closest-hit test
// `t < threshold`, utilities.wgsl:523, decided without finishing the closest
// hit; see the equivalence argument in traverse_blas).
template <bool ANY_HIT>
__device__ __forceinline__ bool traverse_tlas(const DevScene &sc, V3 o, V3 d, float &rec_t, uint32_t &rec_inst,
                                              uint32_t &rec_tri, float &rec_u, float &rec_v, float threshold,
                                              LdsStack &stk) {
    float4 r_lo = sc.tlas_nodes[0];
    float4 r_hi = sc.tlas_nodes[1];
    if (__float_as_uint(r_lo.w) == 0u && __float_as_uint(r_hi.w) == 0u) return false;
    // Box tests cull by the current closest hit (closest) or by a slack
    // bound above the threshold (any-hit).
    float cull_t = rec_t;
    if (ANY_HIT) cull_t = threshold + world_bound_slack(o, threshold);

    int sp = 0;
    uint32_t node = 0;
    float4 n_lo = r_lo, n_hi = r_hi;
    int iteration = 0;
    while (iteration < kMaxIterations) {
        iteration += 1;
        uint32_t first = __float_as_uint(n_lo.w);
        uint32_t count = __float_as_uint(n_hi.w);
        if (count > 0u) {
            for (uint32_t i = 0; i < count; ++i) {
                uint32_t inst = sc.instance_indices[first + i];
                ObjHit h;
                float new_t = traverse_blas(sc, o, d, inst, ANY_HIT ? cull_t : rec_t, stk, sp, h);
                if (ANY_HIT) {
                    if (new_t < threshold) return true;
                } else if (new_t < rec_t) {
                    // a miss with a finite world t: the reference's unhit
                    // object record, instance 0 (utilities.wgsl:265)
                    rec_t = new_t;
                    rec_inst = h.found ? inst : 0u;
                    rec_tri = h.tri;
                    rec_u = h.u;
                    rec_v = h.v;
                    cull_t = rec_t;
                }
            }
            if (sp == 0) break;
            sp -= 1;
            node = stk.get(sp);
            n_lo = sc.tlas_nodes[2 * node];
            n_hi = sc.tlas_nodes[2 * node + 1];
            continue;
        }
        uint32_t a = first, b = first + 1u;
        const float4 *pa = sc.tlas_nodes + 2 * a;
        float4 a_lo = pa[0], a_hi = pa[1], b_lo = pa[2], b_hi = pa[3];
        float da = slab(o, d, cull_t, a_lo, a_hi);
        float db = slab(o, d, cull_t, b_lo, b_hi);
        if (da > db) {
            float tf = da; da = db; db = tf;
            uint32_t ti = a; a = b; b = ti;
            float4 t0 = a_lo; a_lo = b_lo; b_lo = t0;
            float4 t1 = a_hi; a_hi = b_hi; b_hi = t1;
        }
        if (da == kTFar) {
            if (sp == 0) break;
            sp -= 1;
            node = stk.get(sp);
            n_lo = sc.tlas_nodes[2 * node];
            n_hi = sc.tlas_nodes[2 * node + 1];
        } else {
            node = a;
            n_lo = a_lo;
            n_hi = a_hi;
            if (db != kTFar) {
                if (sp < stk.capacity) {
                    stk.put(sp, b);
                    sp += 1;
                } else {
                    atomicOr(sc.error_flags, kErrStackOverflow);
                }
            }
        }
    }
    if (iteration >= kMaxIterations) atomicOr(sc.error_flags, kErrIterationCap);
    return false;
}

// ---------------------------------------------------------------------------
// R8: GGX terms and sampling, utilities.wgsl:468-476, 719-769, 787-872
// (templated on the scalar S: float, or Dual<N> in the adjoint replay)
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ V3T<S> fresnel_schlick(const S &cos_theta, const V3 &f0) {
    S p = pow5_(1.0f - cos_theta);
    return {f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
}
template <class S>
__device__ __forceinline__ S D_GGX(const S &NdotM, const S &roughness) {
    S a = roughness * roughness;
    S a2 = a * a;
    S NdotM2 = NdotM * NdotM;
    S b = (NdotM2 * (a2 - 1.0f) + 1.0f);
    return a2 * kINV_PI / (b * b);
}
template <class S>
__device__ __forceinline__ S G1(const S &NdotO, const S &roughness) {
    S a = roughness * roughness;
    S k = a / 2.0f;
    return NdotO / (NdotO * (1.0f - k) + k);
}
template <class S>
__device__ __forceinline__ S G_smith(const S &NdotO, const S &NdotI, const S &roughness) {
    return G1(NdotI, roughness) * G1(NdotO, roughness);
}

// Material; the roughness has the scalar type of the replay (its tangent is
// the derivative direction of the adjoint), the rest stays f32.
template <class S>
struct MaterialT {
    V3 base, emissive;
    S roughness;
    float reflectance, metallic;
};
using Material = MaterialT<float>;
__device__ __forceinline__ Material load_material(const pulse_material &m) {
    Material r;
    r.base = ld3(m.base_color);
    r.emissive = ld3(m.emissive);
    r.roughness = m.perceptual_roughness;
    r.reflectance = m.reflectance;
    r.metallic = m.metallic;
    return r;
}
template <class S>
__device__ __forceinline__ V3 material_f0(const MaterialT<S> &m) {
    float f = 0.16f * m.reflectance * m.reflectance;
    float om = 1.0f - m.metallic;
    return mk(f * om + m.base.x * m.metallic, f * om + m.base.y * m.metallic, f * om + m.base.z * m.metallic);
}

// sample_cosine_hemisphere, utilities.wgsl:468-476
template <class S>
__device__ __forceinline__ V3T<S> sample_cosine_hemisphere(const V3T<S> &n, float e0, float e1) {
    float cos_theta = 2.0f * e0 - 1.0f;
    float phi = 2.0f * kPI * e1;
    float sin_theta = __builtin_sqrtf(fmaxf(1.0f - cos_theta * cos_theta, 0.0f));
    float sin_phi = pt_sinf(phi);
    float cos_phi = pt_cosf(phi);
    V3 dir = normalize(mk(sin_theta * cos_phi, cos_theta, sin_theta * sin_phi));
    return normalize(add(n, dir));
}

template <class S>
struct SampleT {
    V3T<S> wi, reflectance;
    bool diffuse;     // lobe taken
    float dbase;      // d reflectance_c / d base_c at P held (specular: through f0, 0 unless metallic)
    float NdotO, NdotI;  // values, for the optional horizon clamp
    // Detached roughness derivative of the sampling weight w = brdf / pdf
    // with the sample held fixed and the density it was drawn from, q,
    // differentiated too (score function): d ln(w q)/d roughness.  Specular
    // sample: w q = brdf (F and the lobe pick do not depend on roughness), so
    // d ln(D G); when the pdf floor replaced pdf by 1, w q = brdf q, one more
    // d ln D.  Diffuse sample: 0.
    float dlogf_dr;
    // d ln q / d roughness at the fixed sample (specular: d ln D; diffuse: 0),
    // so that d ln w = dlogf_dr - dlogq_dr.
    float dlogq_dr;
    // The lobe choice (probability P = F0max for the specular lobe, 1 -
    // F0max for the diffuse one) is a decision: the weight divides by P held
    // at its value (P's own tangent is dropped: with the choice's score it
    // cancels in every later contribution), `lobe_p` keeps P with its
    // tangent for what does not cancel; d ln P / d base_c0 (c0 the channel
    // of F0max; nonzero for metallic materials only) and whether the pdf
    // floor replaced the density (then the weight holds no P and the
    // score stays).
    S lobe_p;
    float dlogp_db;
    uint32_t lobe_c0;
    bool floored;
};

// d ln D_GGX / d roughness at a fixed microfacet normal (D = a2 / (pi b^2),
// a2 = r^4, b = NdotM^2 (a2 - 1) + 1).
__device__ __forceinline__ float dlog_d_dr(float r, float NdotM) {
    float a2 = (r * r) * (r * r);
    float m2 = NdotM * NdotM;
    float b = m2 * (a2 - 1.0f) + 1.0f;
    return 4.0f / r - 8.0f * r * r * r * m2 / b;
}
// d ln G_smith / d roughness at fixed directions (G1(x) = x / (x (1 - k) + k),
// k = r^2 / 2).
__device__ __forceinline__ float dlog_g_dr(float r, float NdotO, float NdotI) {
    float k = r * r / 2.0f;
    return -r * (1.0f - NdotI) / (NdotI * (1.0f - k) + k) - r * (1.0f - NdotO) / (NdotO * (1.0f - k) + k);
}

// Optional horizon clamp (PULSE_FLAG_HORIZON_CLAMP, not in the reference):
// the path ends when the view or the sampled direction is below the shading
// horizon.
template <class S>
__device__ __forceinline__ bool below_horizon(const SampleT<S> &smp) { return !(smp.NdotO > 0.0f) || !(smp.NdotI > 0.0f); }

// importance_sample_ggx_d, utilities.wgsl:816-872
template <class S>
__device__ __forceinline__ SampleT<S> importance_sample_ggx_d(const V3T<S> &n, const V3T<S> &wo,
                                                              const MaterialT<S> &m, uint32_t &rng) {
    V3 f0 = material_f0(m);
    S NdotO = dot(n, wo);
    V3T<S> F0 = fresnel_schlick(NdotO, f0);
    S F0_max = max3(F0);
    float e0 = rand_f(rng);
    float e1 = rand_f(rng);
    float e2 = rand_f(rng);
    V3T<S> brdf, wi;
    S pdf_s = sconst<S>(0.0f), pdf_d = sconst<S>(0.0f);
    SampleT<S> out;
    out.diffuse = !(e0 <= val(F0_max));
    float dbase_num = 0.0f, dlog_d = 0.0f, dspec = 0.0f;
    out.NdotO = val(NdotO);
    out.dlogf_dr = 0.0f;
    if (!out.diffuse) {
        S a = m.roughness * m.roughness;
        S a2 = a * a;
        S theta = acos_(sqrt_((1.0f - e1) / (e1 * (a2 - 1.0f) + 1.0f)));
        float phi = e2 * kTWO_PI;
        // orthonormal_from_normal, utilities.wgsl:719-729
        V3T<S> e_two = n;
        V3T<S> e_three;
        if (!(val(length(sub(n, mk(1.0f, 0.0f, 0.0f)))) < 0.001f)) {
            e_three = normalize(cross(e_two, mk(1.0f, 0.0f, 0.0f)));
        } else {
            e_three = normalize(cross(e_two, mk(0.0f, 1.0f, 0.0f)));
        }
        V3T<S> e_one = normalize(cross(e_two, e_three));
        // spherical_to_cartesian_in_on, utilities.wgsl:732-739: A + (B + C)
        S sin_theta = sin_(theta);
        float sp = pt_sinf(phi), cp = pt_cosf(phi);
        S ct = cos_(theta);
        V3T<S> A = {(e_one.x * sin_theta) * sp, (e_one.y * sin_theta) * sp, (e_one.z * sin_theta) * sp};
        V3T<S> B = {e_two.x * ct, e_two.y * ct, e_two.z * ct};
        V3T<S> C = {(e_three.x * sin_theta) * cp, (e_three.y * sin_theta) * cp, (e_three.z * sin_theta) * cp};
        V3T<S> wm = add(A, add(B, C));
        // reflect, utilities.wgsl:767-769
        S k = 2.0f * dot(wo, wm);
        wi = sub(scale(k, wm), wo);
        S NdotM = dot(n, wm);
        S NdotI = dot(n, wi);
        S OdotM = dot(wo, wm);
        V3T<S> F = fresnel_schlick(OdotM, f0);
        S D = D_GGX(NdotM, m.roughness);
        S G = G_smith(NdotO, NdotI, m.roughness);
        S denom = 4.0f * NdotO * NdotI;
        brdf = {F.x * D * G / denom, F.y * D * G / denom, F.z * D * G / denom};
        pdf_s = D * NdotM / (4.0f * OdotM);
        out.NdotI = val(NdotI);
        dlog_d = dlog_d_dr(val(m.roughness), val(NdotM));
        out.dlogf_dr = dlog_d + dlog_g_dr(val(m.roughness), val(NdotO), val(NdotI));
        // d brdf_c / d base_c = metallic (1 - (1 - o.m)^5) D G / denom (through f0)
        if (m.metallic != 0.0f) dspec = m.metallic * (1.0f - pt_pow5f(1.0f - val(OdotM))) * (val(D) * val(G) / val(denom));
    } else {
        wi = sample_cosine_hemisphere(n, e1, e2);
        S NdotI = dot(n, wi);
        out.NdotI = val(NdotI);
        brdf = {m.base.x * kPI * NdotI, m.base.y * kPI * NdotI, m.base.z * kPI * NdotI};
        pdf_d = kPI * NdotI;
        dbase_num = val(kPI * NdotI);
    }
    const S F0_held = sconst<S>(val(F0_max));
    S pdf = ((1.0f - F0_held) * pdf_d) + (F0_held * pdf_s);
    out.floored = val(pdf) < 0.0001f;
    if (out.floored) {
        pdf = sconst<S>(1.0f);
        out.dlogf_dr += dlog_d;
    }
    out.dlogq_dr = dlog_d;
    out.wi = wi;
    out.reflectance = divs(brdf, pdf);
    out.lobe_p = out.diffuse ? 1.0f - F0_max : F0_max;
    if constexpr (!std::is_same<S, float>::value) {
        // floored: the weight holds no P, the choice's score d ln P stays (value x1)
        if (out.floored && val(out.lobe_p) > 0.0f) out.reflectance = scale(out.lobe_p / val(out.lobe_p), out.reflectance);
    }
    out.dbase = (out.diffuse ? dbase_num : dspec) / val(pdf);
    const float F0m = val(F0_max);
    out.lobe_c0 = val(F0.x) == F0m ? 0u : (val(F0.y) == F0m ? 1u : 2u);
    const float dF0 = m.metallic * (1.0f - pt_pow5f(1.0f - val(NdotO)));
    out.dlogp_db = m.metallic == 0.0f ? 0.0f : (out.diffuse ? -dF0 / (1.0f - F0m) : dF0 / F0m);
    return out;
}

// ---------------------------------------------------------------------------
// R10-R12: NEE setup and unoccluded contribution
// ---------------------------------------------------------------------------
// sample_light_triangle_area_cdf (utilities.wgsl:677-691) as a binary search
// for the same index: the reference's loop (mid = l + (r - l) = r) scans from
// the top for the last entry <= e and returns min(that + 1, count - 1); the
// predicate cdf <= e holds on a prefix of the CDF (running sums of
// non-negative areas; NaN entries, if any, form a suffix and compare false),
// so the partition point found here is the same index.
__device__ __forceinline__ uint32_t cdf_pick(const float *cdf, float e, uint32_t count) {
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] <= e) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return min(lo, count - 1u);
}

template <class S>
struct NeeSampleT {
    V3T<S> origin, dir;  // shadow ray
    S threshold;         // occluded iff closest t < threshold
    V3T<S> contrib;      // direct light if unoccluded (before * throughput)
    V3 d_base;           // d contrib_c / d base_c (Lambert evaluator; GGX through f0 when metallic)
    V3 d_emissive;       // d contrib_c / d emissive_c of the light's material
    V3 d_rough;          // d contrib_c / d roughness at fixed light sample (GGX evaluator)
    uint32_t light_material;
};

// sample_direct_light / sample_direct_light_ggx (utilities.wgsl:500-558,
// 562-655) up to the occlusion test, plus the contribution they return when
// the light is visible.  Consumes 4 random numbers.
// light_power: the light is picked by emission strength (the reference's
// intended sample_light_emission_strength_cdf, utilities.wgsl:501, 659-673,
// as a binary search with the same result as its scan) with pdf = the CDF
// step (541-546); otherwise uniformly (502, 540).
template <class S>
__device__ __forceinline__ NeeSampleT<S> nee_sample(const DevScene &sc, const V3T<S> &p0, const V3T<S> &n0,
                                                    const MaterialT<S> &m, const V3T<S> &wo, bool ggx, uint32_t &rng,
                                                    bool light_power = false) {
    NeeSampleT<S> ns;
    uint32_t light_index = light_power ? cdf_pick(sc.strength_cdf, rand_f(rng), sc.light_count)
                                       : rand_u(rng) % sc.light_count;
    pulse_light_data_index ldi = sc.lights[light_index];
    const pulse_mesh_instance &mi = sc.instances[ldi.mesh_instance_index];
    float e0 = rand_f(rng);
    float e1 = rand_f(rng);
    float e2 = rand_f(rng);
    uint32_t prim_index = cdf_pick(sc.light_cdfs + ldi.cdf_offset, e0, mi.triangle_count);
    const pulse_primitive &prim = sc.primitives[mi.triangle_offset + prim_index];
    V3 q0 = ld3(prim.p_first), q1 = ld3(prim.p_second), q2 = ld3(prim.p_third);
    // sample_triangle_uniformly, utilities.wgsl:695-703
    V3 sa = sub(q1, q0), sb = sub(q2, q0);
    V3 pl_obj;
    if (e1 + e2 < 1.0f) {
        pl_obj = add(add(q0, scale(e1, sa)), scale(e2, sb));
    } else {
        pl_obj = add(add(q0, scale(1.0f - e1, sa)), scale(1.0f - e2, sb));
    }
    float ow[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) ow[i] = mi.object_world[i];
    V3 pl = transform_position(ow, pl_obj);
    V3T<S> to_light = sub(pl, p0);
    ns.origin = add(p0, scale(0.001f, n0));
    ns.dir = normalize(to_light);
    ns.threshold = length(sub(pl, ns.origin)) - 0.005f;

    // light normal from the object-space triangle (utilities.wgsl:533-538)
    V3 nl = normalize(cross(sa, sb));
    if (val(dot(nl, neg(ns.dir))) < 0.0f) nl = neg(nl);
    float light_pdf = light_power ? (light_index == 0u ? sc.strength_cdf[0]
                                                       : sc.strength_cdf[light_index] - sc.strength_cdf[light_index - 1u])
                                  : 1.0f / static_cast<float>(sc.light_count);
    float triangle_pdf = 1.0f / sc.light_areas[light_index];
    float pdf = light_pdf * triangle_pdf;
    S cos_r = dot(ns.dir, n0);
    S cos_e = dot(neg(ns.dir), nl);
    V3T<S> brdf;
    float dlog_r = 0.0f, dspec = 0.0f;
    if (!ggx) {
        brdf = lift<S>(mk(m.base.x * kINV_PI, m.base.y * kINV_PI, m.base.z * kINV_PI));
    } else {
        V3 f0 = material_f0(m);
        S NdotO = dot(n0, wo);
        V3T<S> wi = to_light;  // unnormalized, as in utilities.wgsl:639
        V3T<S> wm = normalize(add(scale(0.5f, wi), scale(0.5f, wo)));
        S NdotM = dot(n0, wm);
        S NdotI = dot(n0, wi);
        S OdotM = dot(wo, wm);
        V3T<S> F = fresnel_schlick(OdotM, f0);
        S D = D_GGX(NdotM, m.roughness);
        S G = G_smith(NdotO, NdotI, m.roughness);
        S denom = 4.0f * NdotO * NdotI;
        brdf = {F.x * D * G / denom, F.y * D * G / denom, F.z * D * G / denom};
        dlog_r = dlog_d_dr(val(m.roughness), val(NdotM)) + dlog_g_dr(val(m.roughness), val(NdotO), val(NdotI));
        // d brdf_c / d base_c through f0 (metallic materials)
        if (m.metallic != 0.0f) dspec = m.metallic * (1.0f - pt_pow5f(1.0f - val(OdotM))) * (val(D) * val(G) / val(denom));
    }
    ns.light_material = mi.material_index;
    V3 le = ld3(sc.materials[mi.material_index].emissive);
    V3T<S> dn = sub(n0, nl);
    S dsq = dot(dn, dn);
    V3T<S> dl;
    dl.x = brdf.x * le.x * cos_r * cos_e / pdf / dsq;
    dl.y = brdf.y * le.y * cos_r * cos_e / pdf / dsq;
    dl.z = brdf.z * le.z * cos_r * cos_e / pdf / dsq;
    ns.contrib = {fmax_(dl.x, 0.0f), fmax_(dl.y, 0.0f), fmax_(dl.z, 0.0f)};
    // partial derivatives of the clamped contribution (values)
    float g = val(cos_r) * val(cos_e) / pdf / val(dsq);
    ns.d_emissive = mk(val(dl.x) > 0.0f ? val(brdf.x) * g : 0.0f, val(dl.y) > 0.0f ? val(brdf.y) * g : 0.0f,
                       val(dl.z) > 0.0f ? val(brdf.z) * g : 0.0f);
    ns.d_rough = mk(val(dl.x) > 0.0f ? val(dl.x) * dlog_r : 0.0f, val(dl.y) > 0.0f ? val(dl.y) * dlog_r : 0.0f,
                    val(dl.z) > 0.0f ? val(dl.z) * dlog_r : 0.0f);
    const float db = ggx ? dspec : kINV_PI;
    ns.d_base = mk(val(dl.x) > 0.0f ? db * le.x * g : 0.0f, val(dl.y) > 0.0f ? db * le.y * g : 0.0f,
                   val(dl.z) > 0.0f ? db * le.z * g : 0.0f);
    return ns;
}
using NeeSample = NeeSampleT<float>;
using Sample = SampleT<float>;

// ---------------------------------------------------------------------------
// R2 / R7: primary ray and hit shading
// ---------------------------------------------------------------------------
__device__ __forceinline__ void primary_ray(const pulse_camera &cam, uint32_t x, uint32_t y, V3 &o, V3 &d) {
    float uvx = static_cast<float>(x) / static_cast<float>(cam.width);
    float uvy = static_cast<float>(y) / static_cast<float>(cam.height);
    float cx = uvx * 2.0f - 1.0f;
    float cy = uvy * 2.0f - 1.0f;
    float h[4];
    mat_vec(cam.inverse_view_proj, mk(cx, -cy, 1.0f), 1.0f, h);
    o = ld3(cam.world_position);
    d = normalize(sub(mk(h[0] / h[3], h[1] / h[3], h[2] / h[3]), o));
}

template <class S>
struct HitInfoT {
    V3T<S> normal, position;
    uint32_t material_index;
};
using HitInfo = HitInfoT<float>;
// path_tracer.wgsl:89-98
template <class S>
__device__ __forceinline__ HitInfoT<S> shade_hit(const DevScene &sc, const V3T<S> &o, const V3T<S> &d, const S &t,
                                                 uint32_t inst, uint32_t tri, const S &u, const S &v) {
    const pulse_mesh_instance &mi = sc.instances[inst];
    uint32_t slot = mi.index_offset + tri;
    float4 a = sc.nrm_bvh[3 * slot], b = sc.nrm_bvh[3 * slot + 1], c = sc.nrm_bvh[3 * slot + 2];
    S w = 1.0f - (u + v);
    V3T<S> nrm = add(add(scale(w, mk(a.x, a.y, a.z)), scale(u, mk(b.x, b.y, b.z))), scale(v, mk(c.x, c.y, c.z)));
    float ow[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) ow[i] = mi.object_world[i];
    HitInfoT<S> h;
    h.normal = normalize(transform_direction(ow, nrm));
    h.position = add(o, scale(t, d));
    h.material_index = mi.material_index;
    return h;
}

}  // namespace pt
