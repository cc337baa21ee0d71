// pulse_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// Scalar CPU restatement of pulse's path-tracing hot path, used as the oracle
// that the HIP kernels are checked against.  Only tests/, the smoke check and
// the CPU-baseline leg of bench.py may load it; the product never does.
//
// It is a line-by-line transcription of the reference WGSL, templated on the
// arithmetic type: `float` reproduces the reference's f32 operation order
// (with -ffp-contract=off every product and sum rounds as written), `double`
// is the accuracy oracle.  Each function cites the WGSL it follows.
//
// Branch decisions (closest-hit ids, lobe, Russian roulette, occlusion, ...)
// go through a DecisionTape: a float run can record them and a double run can
// replay them, so the FP64 image is computed along exactly the FP32 paths
// (decision replay) and the number of decisions that would have flipped is
// counted separately.
//
// Pinning (see tests/test_oracle_*.py and the other CPU tests): PCG known
// answers (closed form), BVH closest hit == brute force over all instanced
// triangles, sampler statistics against their densities, GGX closed forms,
// FP32 vs FP64 agreement under decision replay, NEE against a closed-form
// emitter case, the emission-strength pick == clamped upper bound with an
// unchanged image mean, the GI pass against emission-only and closed-form
// direct-light cases, the gradient by step independence, Euler's identity
// and free-decision differences.  Parity partially pinned: the GGX NEE
// evaluator (sample_direct_light_ggx, dead code in the reference, with an
// unnormalised wi) has no closed form here; it is checked only through
// f32/f64 replay agreement and GPU bit-exactness.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/pulse_hip.h"
#include "../pulse_amd/csrc/pt_math.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Scalar traits: f32 uses the portable transcendentals shared with the
// kernels (pt_math.h, checked against libm to <= 1.5 ulp); f64 uses libm.
// ---------------------------------------------------------------------------
template <class R>
struct M;
template <>
struct M<float> {
    static float sqrt(float x) { return std::sqrt(x); }
    static float acos(float x) { return pt_acosf(x); }
    static float sin(float x) { return pt_sinf(x); }
    static float cos(float x) { return pt_cosf(x); }
    static float pow5(float x) { return pt_pow5f(x); }
    static float fmin(float a, float b) { return std::fmin(a, b); }
    static float fmax(float a, float b) { return std::fmax(a, b); }
};
template <>
struct M<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double acos(double x) { return std::acos(x); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double pow5(double x) { return std::pow(x > 0.0 ? x : 0.0, 5.0); }
    static double fmin(double a, double b) { return std::fmin(a, b); }
    static double fmax(double a, double b) { return std::fmax(a, b); }
};

// utilities.wgsl:32-35 (f32 literals; the f64 oracle uses the same f32 values
// so both compute the same estimator)
constexpr float kPI = 3.14159265358f;
constexpr float kTWO_PI = 6.28318530718f;
constexpr float kINV_PI = 0.31830988618f;
constexpr float kTFar = 1e30f;

template <class R>
struct V3 {
    R x, y, z;
};
template <class R>
inline V3<R> v3(R x, R y, R z) { return {x, y, z}; }
template <class R>
inline V3<R> load3(const float *p) { return {static_cast<R>(p[0]), static_cast<R>(p[1]), static_cast<R>(p[2])}; }
template <class R>
inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R>
inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R>
inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R>
inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R>
inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R>
inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R>
inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R>
inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R>
inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R>
inline R length(V3<R> a) { return M<R>::sqrt(dot(a, a)); }
template <class R>
inline V3<R> normalize(V3<R> a) { return a / length(a); }
// length_sq / distance_sq, utilities.wgsl:40-46
template <class R>
inline R length_sq(V3<R> v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
template <class R>
inline R distance_sq(V3<R> a, V3<R> b) { return length_sq(a - b); }

// mat4x4f * vec4f, column-major, accumulated column by column.
template <class R>
inline void mat_vec4(const float *m, R x, R y, R z, R w, R out[4]) {
    for (int r = 0; r < 4; ++r) {
        out[r] = static_cast<R>(m[r]) * x + static_cast<R>(m[4 + r]) * y + static_cast<R>(m[8 + r]) * z +
                 static_cast<R>(m[12 + r]) * w;
    }
}
// transform_position / transform_direction, utilities.wgsl:751-759
template <class R>
inline V3<R> transform_position(const float *m, V3<R> p) {
    R h[4];
    mat_vec4<R>(m, p.x, p.y, p.z, R(1), h);
    return {h[0] / h[3], h[1] / h[3], h[2] / h[3]};
}
template <class R>
inline V3<R> transform_direction(const float *m, V3<R> p) {
    R h[4];
    mat_vec4<R>(m, p.x, p.y, p.z, R(0), h);
    return {h[0], h[1], h[2]};
}

// ---------------------------------------------------------------------------
// R1: PCG hash RNG, utilities.wgsl:67-87; seed path_tracer.wgsl:60-61
// ---------------------------------------------------------------------------
inline uint32_t rand_u(uint32_t &state) {
    state = state * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline float rand_f(uint32_t &state) {
    uint32_t w = rand_u(state);
    uint32_t bits = 0x2f800004u;
    float scale;
    std::memcpy(&scale, &bits, 4);
    return static_cast<float>(w) * scale;
}
inline uint32_t rand_range_u(uint32_t n, uint32_t &state) { return rand_u(state) % n; }
inline uint32_t pixel_seed(uint32_t pixel_index, uint32_t accumulation_count) {
    return pixel_index * 1235243u + accumulation_count * 5817321u;
}

// ---------------------------------------------------------------------------
// Scene view over the 13 bindings (host memory, encase layout).
// ---------------------------------------------------------------------------
struct Scene {
    pulse_scene_uniform uniform{};
    const pulse_primitive *primitives = nullptr;
    const pulse_triangle_data *triangle_data = nullptr;
    const uint32_t *triangle_indices = nullptr;
    const pulse_blas_node *blas_nodes = nullptr;
    const pulse_tlas_node *tlas_nodes = nullptr;
    const uint32_t *instance_indices = nullptr;
    const pulse_mesh_instance *instances = nullptr;
    const pulse_material *materials = nullptr;
    const float *light_triangle_area_cdfs = nullptr;
    const float *light_mesh_areas = nullptr;
    const float *light_emission_strength_cdf = nullptr;
    size_t n_strength_cdf = 0;
    const pulse_light_data_index *light_indices = nullptr;
    size_t n_primitives = 0, n_instances = 0, n_tlas = 0, n_blas = 0;
    bool cdf_linear_scan = false;  // true: run the reference's O(n) scan as written
    // true (f32 only): the kernels' traversal -- each BLAS starts from a
    // conservative object-space bound derived from the current world-space
    // closest hit (or occlusion threshold) and shadow rays stop at the first
    // occluding instance.  Results must equal the reference traversal; the
    // work counters then describe what the kernels execute.
    bool cull_world = false;
    // f64 only: per-material parameters {base.rgb, roughness, emissive.rgb}
    // replacing the buffer's values (finite differences of the adjoint).
    const double *mat_override = nullptr;
};

// Per-ray traversal counters (the algorithmic-bytes model counts these).
struct Counters {
    uint64_t tlas_interior = 0;   // interior TLAS node visits (2 child boxes each)
    uint64_t tlas_leaf = 0;
    uint64_t blas_interior = 0;
    uint64_t blas_leaf = 0;
    uint64_t instances_entered = 0;
    uint64_t tri_tests = 0;
    uint64_t rays = 0;
    uint64_t max_stack = 0;       // deepest per-level stack seen (reference cap 32)
    uint64_t iteration_cap_hits = 0;
    uint64_t shadow_rays = 0;     // NEE occlusion rays
    uint64_t primary_rays = 0;    // depth-0 closest-hit rays
    // Work per ray category (0 primary, 1 bounce, 2 shadow):
    // {tlas_interior, tlas_leaf, blas_interior, blas_leaf, instances, tri_tests, rays}
    uint64_t cat[3][7] = {};
    void add(const Counters &o) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 7; ++b) cat[a][b] += o.cat[a][b];
        tlas_interior += o.tlas_interior;
        tlas_leaf += o.tlas_leaf;
        blas_interior += o.blas_interior;
        blas_leaf += o.blas_leaf;
        instances_entered += o.instances_entered;
        tri_tests += o.tri_tests;
        rays += o.rays;
        max_stack = std::max(max_stack, o.max_stack);
        iteration_cap_hits += o.iteration_cap_hits;
        shadow_rays += o.shadow_rays;
        primary_rays += o.primary_rays;
    }
};

// Adds the traversal work done while it is alive to category `k`.
struct CatScope {
    Counters &c;
    int k;
    uint64_t b[7];
    CatScope(Counters &c_, int k_) : c(c_), k(k_) { snap(b); }
    void snap(uint64_t *v) const {
        v[0] = c.tlas_interior; v[1] = c.tlas_leaf; v[2] = c.blas_interior; v[3] = c.blas_leaf;
        v[4] = c.instances_entered; v[5] = c.tri_tests; v[6] = c.rays;
    }
    ~CatScope() {
        uint64_t a[7];
        snap(a);
        for (int i = 0; i < 7; ++i) c.cat[k][i] += a[i] - b[i];
    }
};

template <class R>
struct Ray {
    V3<R> origin, dir;
    R t;  // record.t
    uint32_t instance_index = 0, triangle_index = 0;
    R u = 0, v = 0;
};

template <class R>
inline R sentinel() { return static_cast<R>(kTFar); }

// ray_aabb_intersect, utilities.wgsl:405-426
template <class R>
R ray_aabb_intersect(const Ray<R> &ray, const float *mn, const float *mx) {
    R tx1 = (static_cast<R>(mn[0]) - ray.origin.x) / ray.dir.x;
    R tx2 = (static_cast<R>(mx[0]) - ray.origin.x) / ray.dir.x;
    R tmin = M<R>::fmin(tx1, tx2);
    R tmax = M<R>::fmax(tx1, tx2);
    R ty1 = (static_cast<R>(mn[1]) - ray.origin.y) / ray.dir.y;
    R ty2 = (static_cast<R>(mx[1]) - ray.origin.y) / ray.dir.y;
    tmin = M<R>::fmax(tmin, M<R>::fmin(ty1, ty2));
    tmax = M<R>::fmin(tmax, M<R>::fmax(ty1, ty2));
    R tz1 = (static_cast<R>(mn[2]) - ray.origin.z) / ray.dir.z;
    R tz2 = (static_cast<R>(mx[2]) - ray.origin.z) / ray.dir.z;
    tmin = M<R>::fmax(tmin, M<R>::fmin(tz1, tz2));
    tmax = M<R>::fmin(tmax, M<R>::fmax(tz1, tz2));
    if (tmax >= tmin && tmin < ray.t && tmax > R(0)) return tmin;
    return sentinel<R>();
}

// get_primitive, utilities.wgsl:100-104
inline const pulse_primitive &get_primitive(const Scene &s, uint32_t tri, uint32_t inst) {
    const pulse_mesh_instance &mi = s.instances[inst];
    uint32_t ti = s.triangle_indices[tri + mi.index_offset];
    return s.primitives[ti + mi.triangle_offset];
}

// Moeller-Trumbore, utilities.wgsl:430-460 (object-space ray).
template <class R>
void ray_triangle_intersect(const Scene &s, Ray<R> &ray, uint32_t tri, uint32_t inst, Counters &c) {
    c.tri_tests++;
    const pulse_primitive &prim = get_primitive(s, tri, inst);
    V3<R> p0 = load3<R>(prim.p_first), p1 = load3<R>(prim.p_second), p2 = load3<R>(prim.p_third);
    V3<R> edge_1 = p1 - p0;
    V3<R> edge_2 = p2 - p0;
    V3<R> h = cross(ray.dir, edge_2);
    R a = dot(edge_1, h);
    if (std::fabs(a) < R(0.0001f)) return;
    R f = R(1) / a;
    V3<R> sv = ray.origin - p0;
    R u = f * dot(sv, h);
    if (u < R(0) || u > R(1)) return;
    V3<R> q = cross(sv, edge_1);
    R v = f * dot(ray.dir, q);
    if (v < R(0) || u + v > R(1)) return;
    R t = f * dot(edge_2, q);
    if (t > R(0.001f) && t < ray.t) {
        ray.t = t;
        ray.instance_index = inst;
        ray.triangle_index = tri;
        ray.u = u;
        ray.v = v;
    }
}

constexpr int kMaxIterations = 100000;  // utilities.wgsl:137, 271
constexpr int kOracleStack = 256;       // reference: 32 (no overflow check)

// traverse_blas, utilities.wgsl:259-326
// World t computed by the last traverse_blas call (any-hit bookkeeping).
inline thread_local double ray_world_t_last = 0.0;

// Conservative object-space bound for a world-space bound (culling mode):
// the world bound is padded by 2^-12 of max(|origin|, bound) + 2^-20 to cover
// the rounding of the reference's world-t conversion, converted with the
// length of the transformed direction and padded by 2^-10.
inline float cull_bound(V3<float> o, float world_bound, float dlen) {
    if (!(world_bound < kTFar)) return kTFar;
    float m = std::fmax(std::fmax(std::fabs(o.x), std::fabs(o.y)), std::fmax(std::fabs(o.z), world_bound));
    float wb = world_bound + (m * 0x1p-12f + 0x1p-20f);
    float ob = wb * dlen * (1.0f + 0x1p-10f);
    return ob < kTFar ? ob : kTFar;
}

// Object t below which a hit certainly converts to a world t below
// `threshold` (occlusion early exit of the kernels); -1 disables it.
inline thread_local float blas_safe_t_world = -1.0f;

template <class R>
void traverse_blas(const Scene &s, Ray<R> &ray, uint32_t inst, Counters &c, R world_bound = sentinel<R>()) {
    c.instances_entered++;
    const pulse_mesh_instance &mi = s.instances[inst];
    Ray<R> ro;
    ro.origin = transform_position<R>(mi.world_object, ray.origin);
    V3<R> draw = transform_direction<R>(mi.world_object, ray.dir);
    ro.dir = normalize(draw);
    ro.t = sentinel<R>();
    bool culling = false;
    float safe_t = -1.0f;
    if constexpr (std::is_same<R, float>::value) {
        if (s.cull_world) {
            float dlen = length(draw);
            ro.t = cull_bound(ray.origin, world_bound, dlen);
            culling = true;
            if (blas_safe_t_world > 0.0f) safe_t = blas_safe_t_world * dlen * (1.0f - 0x1p-10f);
        }
    }
    const R start_t = ro.t;

    uint32_t node_index = 0;
    uint32_t stack[kOracleStack];
    int sp = 0;
    int iteration = 0;
    while (iteration < kMaxIterations) {
        iteration += 1;
        const pulse_blas_node &node = s.blas_nodes[node_index + mi.node_offset];
        if (node.tri_count > 0u) {
            c.blas_leaf++;
            for (uint32_t i = 0; i < node.tri_count; ++i) ray_triangle_intersect(s, ro, node.a_or_first_tri + i, inst, c);
            if (culling && safe_t > 0.0f && ro.t < safe_t) {
                ray_world_t_last = -1.0;  // certainly below the threshold
                return;
            }
            if (sp == 0) break;
            sp -= 1;
            node_index = stack[sp];
            continue;
        }
        c.blas_interior++;
        uint32_t a = node.a_or_first_tri, b = a + 1u;
        const pulse_blas_node &ca = s.blas_nodes[a + mi.node_offset];
        const pulse_blas_node &cb = s.blas_nodes[b + mi.node_offset];
        R da = ray_aabb_intersect(ro, ca.aabb_min, ca.aabb_max);
        R db = ray_aabb_intersect(ro, cb.aabb_min, cb.aabb_max);
        if (da > db) {
            std::swap(da, db);
            std::swap(a, b);
        }
        if (da == sentinel<R>()) {
            if (sp == 0) break;
            sp -= 1;
            node_index = stack[sp];
        } else {
            node_index = a;
            if (db != sentinel<R>()) {
                stack[sp] = b;
                sp += 1;
                c.max_stack = std::max<uint64_t>(c.max_stack, sp);
            }
        }
    }
    if (iteration >= kMaxIterations) c.iteration_cap_hits++;

    // World t through the hit point (utilities.wgsl:318-325).  In f32 a BLAS
    // miss (t = 1e30) overflows to +inf here and never updates; f64 does not
    // overflow, so the miss is tested explicitly (same result).
    if (!std::is_same<R, float>::value && ro.t == sentinel<R>()) return;
    if (culling && ro.t == start_t) ro.t = sentinel<R>();  // nothing below the bound: as a miss
    V3<R> hit_obj = ro.origin + ro.t * ro.dir;
    V3<R> hit_world = transform_position<R>(mi.object_world, hit_obj);
    R new_t = length(hit_world - ray.origin);
    ray_world_t_last = static_cast<double>(new_t);
    if (new_t < ray.t) {
        ray.t = new_t;
        ray.instance_index = ro.instance_index;
        ray.triangle_index = ro.triangle_index;
        ray.u = ro.u;
        ray.v = ro.v;
    }
}

// traverse_tlas, utilities.wgsl:127-188
template <class R>
void traverse_tlas(const Scene &s, Ray<R> &ray, Counters &c) {
    c.rays++;
    if (s.tlas_nodes[0].a_or_first_instance == 0u && s.tlas_nodes[0].instance_count == 0u) return;
    uint32_t node_index = 0;
    uint32_t stack[kOracleStack];
    int sp = 0;
    int iteration = 0;
    while (iteration < kMaxIterations) {
        iteration += 1;
        const pulse_tlas_node &node = s.tlas_nodes[node_index];
        if (node.instance_count > 0u) {
            c.tlas_leaf++;
            for (uint32_t i = 0; i < node.instance_count; ++i) {
                traverse_blas(s, ray, s.instance_indices[node.a_or_first_instance + i], c, ray.t);
            }
            if (sp == 0) break;
            sp -= 1;
            node_index = stack[sp];
            continue;
        }
        c.tlas_interior++;
        uint32_t a = node.a_or_first_instance, b = a + 1u;
        R da = ray_aabb_intersect(ray, s.tlas_nodes[a].aabb_min, s.tlas_nodes[a].aabb_max);
        R db = ray_aabb_intersect(ray, s.tlas_nodes[b].aabb_min, s.tlas_nodes[b].aabb_max);
        if (da > db) {
            std::swap(da, db);
            std::swap(a, b);
        }
        if (da == sentinel<R>()) {
            if (sp == 0) break;
            sp -= 1;
            node_index = stack[sp];
        } else {
            node_index = a;
            if (db != sentinel<R>()) {
                stack[sp] = b;
                sp += 1;
                c.max_stack = std::max<uint64_t>(c.max_stack, sp);
            }
        }
    }
    if (iteration >= kMaxIterations) c.iteration_cap_hits++;
}

// The kernels' occlusion query (culling mode): TLAS boxes culled at a padded
// threshold, each entered BLAS searched for its closest hit below the padded
// bound, stop at the first instance whose world t is below the threshold.
// Equals `closest t < threshold` of the full traversal.
inline bool traverse_tlas_anyhit(const Scene &s, Ray<float> &ray, float threshold, Counters &c) {
    c.rays++;
    if (s.tlas_nodes[0].a_or_first_instance == 0u && s.tlas_nodes[0].instance_count == 0u) return false;
    float m = std::fmax(std::fmax(std::fabs(ray.origin.x), std::fabs(ray.origin.y)),
                        std::fmax(std::fabs(ray.origin.z), threshold));
    const float slack = m * 0x1p-12f + 0x1p-20f;
    const float cull_t = threshold + slack;
    ray.t = cull_t;
    struct SafeScope {
        explicit SafeScope(float v) { blas_safe_t_world = v; }
        ~SafeScope() { blas_safe_t_world = -1.0f; }
    } safe_scope(threshold - slack > 0.0f ? threshold - slack : -1.0f);
    uint32_t node_index = 0;
    uint32_t stack[kOracleStack];
    int sp = 0;
    int iteration = 0;
    while (iteration < kMaxIterations) {
        iteration += 1;
        const pulse_tlas_node &node = s.tlas_nodes[node_index];
        if (node.instance_count > 0u) {
            c.tlas_leaf++;
            for (uint32_t i = 0; i < node.instance_count; ++i) {
                Ray<float> probe = ray;
                probe.t = kTFar;
                traverse_blas(s, probe, s.instance_indices[node.a_or_first_instance + i], c, cull_t);
                if (ray_world_t_last < static_cast<double>(threshold)) return true;
            }
            if (sp == 0) break;
            node_index = stack[--sp];
            continue;
        }
        c.tlas_interior++;
        uint32_t a = node.a_or_first_instance, b = a + 1u;
        float da = ray_aabb_intersect(ray, s.tlas_nodes[a].aabb_min, s.tlas_nodes[a].aabb_max);
        float db = ray_aabb_intersect(ray, s.tlas_nodes[b].aabb_min, s.tlas_nodes[b].aabb_max);
        if (da > db) {
            std::swap(da, db);
            std::swap(a, b);
        }
        if (da == kTFar) {
            if (sp == 0) break;
            node_index = stack[--sp];
        } else {
            node_index = a;
            if (db != kTFar) stack[sp++] = b;
        }
    }
    return false;
}

// Brute force closest hit: every instance, every triangle of its mesh, with
// the same object-space test and world-t conversion as traverse_blas.  Hit
// ids equal traversal's except on exact t-ties (first found wins).
template <class R>
void trace_bruteforce(const Scene &s, Ray<R> &ray, Counters &c) {
    for (uint32_t inst = 0; inst < s.uniform.instance_count; ++inst) {
        const pulse_mesh_instance &mi = s.instances[inst];
        Ray<R> ro;
        ro.origin = transform_position<R>(mi.world_object, ray.origin);
        ro.dir = normalize(transform_direction<R>(mi.world_object, ray.dir));
        ro.t = sentinel<R>();
        for (uint32_t k = 0; k < mi.triangle_count; ++k) ray_triangle_intersect(s, ro, k, inst, c);
        if (!std::is_same<R, float>::value && ro.t == sentinel<R>()) continue;
        V3<R> hw = transform_position<R>(mi.object_world, ro.origin + ro.t * ro.dir);
        R nt = length(hw - ray.origin);
        if (nt < ray.t) {
            ray.t = nt;
            ray.instance_index = ro.instance_index;
            ray.triangle_index = ro.triangle_index;
            ray.u = ro.u;
            ray.v = ro.v;
        }
    }
}

// ---------------------------------------------------------------------------
// Decision tape: record in one precision, replay in another.
// ---------------------------------------------------------------------------
struct DecisionTape {
    enum Mode { OFF, RECORD, REPLAY };
    Mode mode = OFF;
    std::vector<uint64_t> *log = nullptr;
    size_t pos = 0;
    uint64_t flips = 0;  // replayed decisions that differ from the computed one
    bool probe = true;   // replay: still run the traversals to count flips
    // Replay without probing: the next recorded decision.
    bool skip_probe() const { return mode == REPLAY && !probe; }
    uint64_t next() { return pos < log->size() ? (*log)[pos++] : 0; }
    uint64_t take(uint64_t computed) {
        if (mode == RECORD) {
            log->push_back(computed);
            return computed;
        }
        if (mode == REPLAY) {
            uint64_t v = pos < log->size() ? (*log)[pos] : computed;
            pos++;
            if (v != computed) flips++;
            return v;
        }
        return computed;
    }
    bool take_bool(bool computed) { return take(computed ? 1u : 0u) != 0; }
    // Frozen half vectors (the specular lobe's sampled microfacet normal):
    // RECORD_DIRS stores them with the density q0 they were drawn from
    // within the lobe; FREEZE_DIRS replays them instead of mapping the
    // random numbers through the current roughness, and reweights the
    // sample by the likelihood ratio q(r)/q0 (importance reweighting keeps
    // the estimator's expectation that of the perturbed parameters).  The
    // roughness derivative of a path so replayed is the detached
    // (score-function) estimator of the derivative of the expected image.
    enum DirMode { DIRS_OFF, RECORD_DIRS, FREEZE_DIRS };
    DirMode dir_mode = DIRS_OFF;
    std::vector<double> *dirs = nullptr;
    size_t dir_pos = 0;
    // Frozen roulette and lobe choices: RECORD_RR stores, in path order,
    // each roulette survival probability p = max(throughput) and each
    // GGX sample's lobe probability (F0max or 1 - F0max); FREEZE_RR
    // replays the decisions and treats a recorded p <= 1 (a random
    // survival) and a lobe choice as decisions sampled at the unperturbed
    // parameters: the path takes the perturbed estimator's own p' and is
    // reweighted by the likelihood ratio p' / p of the decision, so that
    // its weight stays 1/p while the throughput later roulettes see is the
    // perturbed estimator's (divided by p').  A recorded roulette p > 1 (the
    // path survives whatever the draw) follows the perturbed throughput.
    // The product of the likelihood ratios so far (frozen directions,
    // survivals and lobes) is `lr`: the throughput carries it, the
    // roulette's p is that of the throughput without it.
    enum RrMode { RR_OFF, RECORD_RR, FREEZE_RR };
    RrMode rr_mode = RR_OFF;
    std::vector<double> *rr = nullptr;
    size_t rr_pos = 0;
    double lr = 1.0;
};

constexpr uint64_t kMissCode = ~0ull;
inline uint64_t hit_code(uint32_t inst, uint32_t tri) { return (static_cast<uint64_t>(inst) << 32) | tri; }

// trace_ray (utilities.wgsl:112-114) behind the decision tape: in replay the
// recorded closest (instance, triangle) is intersected directly and its t,
// u, v recomputed at precision R.
template <class R>
void trace_closest(const Scene &s, Ray<R> &ray, Counters &c, DecisionTape &tape) {
    if (tape.mode != DecisionTape::REPLAY) {
        traverse_tlas(s, ray, c);
        tape.take(ray.t >= sentinel<R>() ? kMissCode : hit_code(ray.instance_index, ray.triangle_index));
        return;
    }
    uint64_t code;
    if (tape.skip_probe()) {
        code = tape.next();
    } else {
        Ray<R> probe = ray;
        traverse_tlas(s, probe, c);
        code = tape.take(probe.t >= sentinel<R>() ? kMissCode : hit_code(probe.instance_index, probe.triangle_index));
    }
    if (code == kMissCode) {
        ray.t = sentinel<R>();
        return;
    }
    uint32_t inst = static_cast<uint32_t>(code >> 32), tri = static_cast<uint32_t>(code & 0xffffffffu);
    const pulse_mesh_instance &mi = s.instances[inst];
    Ray<R> ro;
    ro.origin = transform_position<R>(mi.world_object, ray.origin);
    ro.dir = normalize(transform_direction<R>(mi.world_object, ray.dir));
    const pulse_primitive &prim = get_primitive(s, tri, inst);
    V3<R> p0 = load3<R>(prim.p_first), p1 = load3<R>(prim.p_second), p2 = load3<R>(prim.p_third);
    V3<R> e1 = p1 - p0, e2 = p2 - p0;
    V3<R> h = cross(ro.dir, e2);
    R f = R(1) / dot(e1, h);
    V3<R> sv = ro.origin - p0;
    R u = f * dot(sv, h);
    V3<R> q = cross(sv, e1);
    R v = f * dot(ro.dir, q);
    R t = f * dot(e2, q);
    V3<R> hw = transform_position<R>(mi.object_world, ro.origin + t * ro.dir);
    ray.t = length(hw - ray.origin);
    ray.instance_index = inst;
    ray.triangle_index = tri;
    ray.u = u;
    ray.v = v;
}

// ---------------------------------------------------------------------------
// R8: GGX helpers and sampling, utilities.wgsl:468-476, 712-769, 787-872
// ---------------------------------------------------------------------------
template <class R>
inline V3<R> fresnel_schlick(R cos_theta, V3<R> f0) {
    R p = M<R>::pow5(R(1) - cos_theta);
    return {f0.x + (R(1) - f0.x) * p, f0.y + (R(1) - f0.y) * p, f0.z + (R(1) - f0.z) * p};
}
template <class R>
inline R D_GGX(R NdotM, R roughness) {
    R a = roughness * roughness;
    R a2 = a * a;
    R NdotM2 = NdotM * NdotM;
    R b = (NdotM2 * (a2 - R(1)) + R(1));
    return a2 * R(kINV_PI) / (b * b);
}
template <class R>
inline R G1_GGX_schlick(R NdotO, R roughness) {
    R a = roughness * roughness;
    R k = a / R(2);
    return NdotO / (NdotO * (R(1) - k) + k);
}
template <class R>
inline R G_smith(R NdotO, R NdotI, R roughness) {
    return G1_GGX_schlick(NdotI, roughness) * G1_GGX_schlick(NdotO, roughness);
}

template <class R>
struct ON {
    V3<R> e_one, e_two, e_three;
};
// orthonormal_from_normal, utilities.wgsl:719-729 (are_aligned 53-55)
template <class R>
ON<R> orthonormal_from_normal(V3<R> normal, DecisionTape &tape) {
    ON<R> on;
    on.e_two = normal;
    V3<R> xh{R(1), R(0), R(0)}, yh{R(0), R(1), R(0)};
    bool aligned = tape.take_bool(length(normal - xh) < R(0.001f));
    if (!aligned) {
        on.e_three = normalize(cross(on.e_two, xh));
    } else {
        on.e_three = normalize(cross(on.e_two, yh));
    }
    on.e_one = normalize(cross(on.e_two, on.e_three));
    return on;
}
// spherical_to_cartesian_in_on, utilities.wgsl:732-739: A + (B + C)
template <class R>
V3<R> spherical_to_cartesian_in_on(R theta, R phi, const ON<R> &on) {
    R sin_theta = M<R>::sin(theta);
    V3<R> A = (on.e_one * sin_theta) * M<R>::sin(phi);
    V3<R> B = on.e_two * M<R>::cos(theta);
    V3<R> C = (on.e_three * sin_theta) * M<R>::cos(phi);
    return A + (B + C);
}
// reflect, utilities.wgsl:767-769
template <class R>
inline V3<R> reflect(V3<R> a, V3<R> b) {
    R k = R(2) * dot(a, b);
    return (k * b) - a;
}
// sample_cosine_hemisphere, utilities.wgsl:468-476
template <class R>
V3<R> sample_cosine_hemisphere(V3<R> normal, R e0, R e1) {
    R cos_theta = R(2) * e0 - R(1);
    R phi = R(2) * R(kPI) * e1;
    R sin_theta = M<R>::sqrt(M<R>::fmax(R(1) - cos_theta * cos_theta, R(0)));
    R sin_phi = M<R>::sin(phi);
    R cos_phi = M<R>::cos(phi);
    V3<R> d = normalize(V3<R>{sin_theta * cos_phi, cos_theta, sin_theta * sin_phi});
    return normalize(normal + d);
}

template <class R>
struct MaterialR {
    V3<R> base, emissive;
    R roughness, reflectance, metallic;
};
template <class R>
inline MaterialR<R> load_material(const pulse_material &m) {
    return {load3<R>(m.base_color), load3<R>(m.emissive), static_cast<R>(m.perceptual_roughness),
            static_cast<R>(m.reflectance), static_cast<R>(m.metallic)};
}
// Material / emission of material `idx`, with the f64 parameter override.
template <class R>
inline MaterialR<R> scene_material(const Scene &s, uint32_t idx) {
    MaterialR<R> m = load_material<R>(s.materials[idx]);
    if (s.mat_override) {
        const double *p = s.mat_override + 7 * static_cast<size_t>(idx);
        m.base = {static_cast<R>(p[0]), static_cast<R>(p[1]), static_cast<R>(p[2])};
        m.roughness = static_cast<R>(p[3]);
        m.emissive = {static_cast<R>(p[4]), static_cast<R>(p[5]), static_cast<R>(p[6])};
    }
    return m;
}

template <class R>
inline V3<R> material_f0(const MaterialR<R> &m) {
    R f = R(0.16f) * m.reflectance * m.reflectance;
    V3<R> f0{f, f, f};
    R om = R(1) - m.metallic;
    return {f0.x * om + m.base.x * m.metallic, f0.y * om + m.base.y * m.metallic, f0.z * om + m.base.z * m.metallic};
}
template <class R>
inline R max3(V3<R> v) { return M<R>::fmax(M<R>::fmax(v.x, v.y), v.z); }

template <class R>
struct SampleResult {
    V3<R> wi, reflectance;
    R NdotO, NdotI;  // for the optional horizon clamp
};

// importance_sample_ggx_d, utilities.wgsl:816-872
template <class R>
SampleResult<R> importance_sample_ggx_d(V3<R> n, V3<R> wo, const MaterialR<R> &mat, uint32_t &rng,
                                        DecisionTape &tape) {
    V3<R> f0 = material_f0(mat);
    R NdotO = dot(n, wo);
    V3<R> F0 = fresnel_schlick(NdotO, f0);
    R F0_max = max3(F0);
    R e0 = static_cast<R>(rand_f(rng));
    R e1 = static_cast<R>(rand_f(rng));
    R e2 = static_cast<R>(rand_f(rng));
    V3<R> brdf, wi;
    R pdf_s = R(0), pdf_d = R(0);
    R NdotI_out = R(0);
    V3<R> wm_out{R(0), R(0), R(0)};
    const bool specular = tape.take_bool(e0 <= F0_max);
    if (specular) {
        R a = mat.roughness * mat.roughness;
        R a2 = a * a;
        R theta = M<R>::acos(M<R>::sqrt((R(1) - e1) / (e1 * (a2 - R(1)) + R(1))));
        R phi = e2 * R(kTWO_PI);
        ON<R> on = orthonormal_from_normal(n, tape);
        V3<R> wm = spherical_to_cartesian_in_on(theta, phi, on);
        if (tape.dir_mode == DecisionTape::FREEZE_DIRS && tape.dir_pos + 4 <= tape.dirs->size()) {
            const double *w = tape.dirs->data() + tape.dir_pos;
            wm = V3<R>{static_cast<R>(w[0]), static_cast<R>(w[1]), static_cast<R>(w[2])};
        }
        wm_out = wm;
        wi = reflect(wo, wm);
        R NdotM = dot(n, wm);
        R NdotI = dot(n, wi);
        R OdotM = dot(wo, wm);
        V3<R> F = fresnel_schlick(OdotM, f0);
        R D = D_GGX(NdotM, mat.roughness);
        R G = G_smith(NdotO, NdotI, mat.roughness);
        R denom = R(4) * NdotO * NdotI;
        brdf = V3<R>{F.x * D * G, F.y * D * G, F.z * D * G} / denom;
        pdf_s = D * NdotM / (R(4) * OdotM);
        NdotI_out = NdotI;
    } else {
        wi = sample_cosine_hemisphere(n, e1, e2);
        R NdotI = dot(n, wi);
        NdotI_out = NdotI;
        brdf = V3<R>{mat.base.x * R(kPI) * NdotI, mat.base.y * R(kPI) * NdotI, mat.base.z * R(kPI) * NdotI};
        pdf_d = R(kPI) * NdotI;
    }
    R pdf = ((R(1) - F0_max) * pdf_d) + (F0_max * pdf_s);
    R ratio = R(1);  // likelihood ratio of a frozen specular sample and of the replayed lobe
    if (specular) {
        if (tape.dir_mode == DecisionTape::RECORD_DIRS) {
            // wm was just sampled; pdf_s is the density it was drawn from
            tape.dirs->insert(tape.dirs->end(), {static_cast<double>(wm_out.x), static_cast<double>(wm_out.y),
                                                 static_cast<double>(wm_out.z), static_cast<double>(pdf_s)});
        } else if (tape.dir_mode == DecisionTape::FREEZE_DIRS && tape.dir_pos + 4 <= tape.dirs->size()) {
            ratio = pdf_s / static_cast<R>((*tape.dirs)[tape.dir_pos + 3]);
            tape.dir_pos += 4;
            tape.lr *= static_cast<double>(ratio);
        }
    }
    const R p_lobe = specular ? F0_max : R(1) - F0_max;
    if (tape.rr_mode == DecisionTape::RECORD_RR) {
        tape.rr->push_back(static_cast<double>(p_lobe));
    } else if (tape.rr_mode == DecisionTape::FREEZE_RR && tape.rr_pos < tape.rr->size()) {
        const R lobe_ratio = p_lobe / static_cast<R>((*tape.rr)[tape.rr_pos++]);
        ratio = ratio * lobe_ratio;
        tape.lr *= static_cast<double>(lobe_ratio);
    }
    if (tape.take_bool(pdf < R(0.0001f))) pdf = R(1);
    return {wi, brdf / pdf * ratio, NdotO, NdotI_out};
}

// ---------------------------------------------------------------------------
// R10-R12: next event estimation, utilities.wgsl:500-703
// ---------------------------------------------------------------------------
// sample_light_triangle_area_cdf, utilities.wgsl:677-691, executed as written
// (`mid = l + (r - l)` makes it a top-down linear scan).
inline uint32_t sample_light_triangle_area_cdf(const Scene &s, float e, uint32_t cdf_offset, uint32_t count) {
    int64_t l = cdf_offset;
    int64_t r = l + static_cast<int64_t>(count) - 1;
    while (l <= r) {
        int64_t mid = l + (r - l);
        if (s.light_triangle_area_cdfs[mid] <= e) {
            l = mid + 1;
        } else {
            r = mid - 1;
        }
    }
    uint32_t lu = static_cast<uint32_t>(l);
    return std::min(lu, cdf_offset + count - 1u) - cdf_offset;
}

// The same pick by binary search: the predicate cdf[i] <= e is true on a
// prefix and false after it (the CDF is a running sum of non-negative areas;
// a NaN area turns every later entry NaN, i.e. false), so the scan above
// returns min(#true, count - 1), which is the partition point found here.
inline uint32_t sample_light_triangle_area_cdf_bsearch(const Scene &s, float e, uint32_t cdf_offset, uint32_t count) {
    uint32_t lo = 0, hi = count;  // first index in [lo, hi) with !(cdf <= e)
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (s.light_triangle_area_cdfs[cdf_offset + mid] <= e) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return std::min(lo, count - 1u);
}

// sample_triangle_uniformly, utilities.wgsl:695-703
template <class R>
V3<R> sample_triangle_uniformly(R e0, R e1, V3<R> p0, V3<R> p1, V3<R> p2, DecisionTape &tape) {
    V3<R> a = p1 - p0;
    V3<R> b = p2 - p0;
    if (tape.take_bool(e0 + e1 < R(1))) return (p0 + (e0 * a)) + (e1 * b);
    return (p0 + ((R(1) - e0) * a)) + ((R(1) - e1) * b);
}

// Occlusion test of the shadow ray (utilities.wgsl:516-525): closest hit,
// occluded iff t_hit < |pl - origin| - 0.005.
template <class R>
struct ShadowSetup {
    Ray<R> ray;
    V3<R> pl;
    pulse_primitive prim;
    uint32_t light_index;
    uint32_t material_index;
    bool light_power;  // light picked by emission strength (else uniformly)
};

// sample_light_emission_strength_cdf, utilities.wgsl:659-673, as written:
// `mid = l + (r - l)` makes it a top-down scan; the result is
// min(#(cdf[i] <= e), light_count - 1).
inline uint32_t sample_light_emission_strength_cdf(const Scene &s, float e) {
    int64_t l = 0;
    int64_t r = static_cast<int64_t>(s.uniform.light_count) - 1;
    while (l <= r) {
        int64_t mid = l + (r - l);
        if (s.light_emission_strength_cdf[mid] <= e) {
            l = mid + 1;
        } else {
            r = mid - 1;
        }
    }
    return std::min(static_cast<uint32_t>(l), s.uniform.light_count - 1u);
}
// The pdf of light i under that pick, as the reference's commented-out code
// (utilities.wgsl:541-546): cdf[0] for the first light, else the difference.
template <class R>
inline R light_strength_pdf(const Scene &s, uint32_t i) {
    if (i == 0) return static_cast<R>(s.light_emission_strength_cdf[0]);
    return static_cast<R>(s.light_emission_strength_cdf[i]) - static_cast<R>(s.light_emission_strength_cdf[i - 1]);
}

template <class R>
ShadowSetup<R> nee_setup(const Scene &s, V3<R> p0, V3<R> n0, uint32_t &rng, DecisionTape &tape, bool light_power) {
    ShadowSetup<R> st;
    // the reference's intended pick (utilities.wgsl:501) or the live one (502)
    uint32_t light_index = light_power ? sample_light_emission_strength_cdf(s, rand_f(rng))
                                       : rand_range_u(s.uniform.light_count, rng);
    st.light_power = light_power;
    const pulse_light_data_index &ldi = s.light_indices[light_index];
    const pulse_mesh_instance &mi = s.instances[ldi.mesh_instance_index];
    float e0 = rand_f(rng);
    R e1 = static_cast<R>(rand_f(rng));
    R e2 = static_cast<R>(rand_f(rng));
    uint32_t prim_index = s.cdf_linear_scan ? sample_light_triangle_area_cdf(s, e0, ldi.cdf_offset, mi.triangle_count)
                                            : sample_light_triangle_area_cdf_bsearch(s, e0, ldi.cdf_offset, mi.triangle_count);
    st.prim = s.primitives[mi.triangle_offset + prim_index];
    V3<R> pl_obj = sample_triangle_uniformly(e1, e2, load3<R>(st.prim.p_first), load3<R>(st.prim.p_second),
                                             load3<R>(st.prim.p_third), tape);
    st.pl = transform_position<R>(mi.object_world, pl_obj);
    V3<R> to_light = st.pl - p0;
    st.ray.origin = p0 + R(0.001f) * n0;
    st.ray.dir = normalize(to_light);
    st.ray.t = sentinel<R>();
    st.light_index = light_index;
    st.material_index = mi.material_index;
    return st;
}

template <class R>
bool nee_occluded(const Scene &s, ShadowSetup<R> &st, Counters &c, DecisionTape &tape) {
    c.shadow_rays++;
    if (tape.skip_probe()) return tape.next() != 0;
    CatScope scope(c, 2);
    Ray<R> r = st.ray;
    R threshold = length(st.pl - st.ray.origin) - R(0.005f);
    if constexpr (std::is_same<R, float>::value) {
        if (s.cull_world) return tape.take_bool(traverse_tlas_anyhit(s, r, threshold, c));
    }
    traverse_tlas(s, r, c);
    return tape.take_bool(r.t < threshold);
}

// Light normal from the object-space triangle, flipped to face the receiver
// (utilities.wgsl:533-538), and the uniform light / area pdf (540-548).
template <class R>
void nee_light_terms(const Scene &s, const ShadowSetup<R> &st, V3<R> &nl, R &pdf, DecisionTape &tape) {
    V3<R> side_a = load3<R>(st.prim.p_second) - load3<R>(st.prim.p_first);
    V3<R> side_b = load3<R>(st.prim.p_third) - load3<R>(st.prim.p_first);
    nl = normalize(cross(side_a, side_b));
    if (tape.take_bool(dot(nl, -st.ray.dir) < R(0))) nl = -nl;
    R light_pdf = st.light_power ? light_strength_pdf<R>(s, st.light_index)
                                 : R(1) / static_cast<R>(static_cast<float>(s.uniform.light_count));
    R triangle_pdf = R(1) / static_cast<R>(s.light_mesh_areas[st.light_index]);
    pdf = light_pdf * triangle_pdf;
}

// sample_direct_light (Lambert), utilities.wgsl:500-558
template <class R>
V3<R> sample_direct_light(const Scene &s, V3<R> p0, V3<R> n0, V3<R> base_color, uint32_t &rng, Counters &c,
                          DecisionTape &tape, bool light_power = false) {
    ShadowSetup<R> st = nee_setup(s, p0, n0, rng, tape, light_power);
    if (nee_occluded(s, st, c, tape)) return {R(0), R(0), R(0)};
    V3<R> nl;
    R pdf;
    nee_light_terms(s, st, nl, pdf, tape);
    R cos_r = dot(st.ray.dir, n0);
    R cos_e = dot(-st.ray.dir, nl);
    V3<R> brdf = base_color * R(kINV_PI);
    V3<R> le = scene_material<R>(s, st.material_index).emissive;
    R dsq = distance_sq(n0, nl);
    V3<R> d = ((((brdf * le) * cos_r) * cos_e) / pdf) / dsq;
    return {M<R>::fmax(d.x, R(0)), M<R>::fmax(d.y, R(0)), M<R>::fmax(d.z, R(0))};
}

// sample_direct_light_ggx, utilities.wgsl:562-655 (dead code in the
// reference; its unused theta/phi/F0 computations are omitted).
template <class R>
V3<R> sample_direct_light_ggx(const Scene &s, V3<R> p0, V3<R> n0, const MaterialR<R> &mat, V3<R> wo, uint32_t &rng,
                              Counters &c, DecisionTape &tape, bool light_power = false) {
    ShadowSetup<R> st = nee_setup(s, p0, n0, rng, tape, light_power);
    if (nee_occluded(s, st, c, tape)) return {R(0), R(0), R(0)};
    V3<R> nl;
    R pdf;
    nee_light_terms(s, st, nl, pdf, tape);
    R cos_r = dot(st.ray.dir, n0);
    R cos_e = dot(-st.ray.dir, nl);
    V3<R> f0 = material_f0(mat);
    R NdotO = dot(n0, wo);
    V3<R> wi = st.pl - p0;  // unnormalized to_light (Q7)
    V3<R> wm = normalize((R(0.5f) * wi) + (R(0.5f) * wo));
    R NdotM = dot(n0, wm);
    R NdotI = dot(n0, wi);
    R OdotM = dot(wo, wm);
    V3<R> F = fresnel_schlick(OdotM, f0);
    R D = D_GGX(NdotM, mat.roughness);
    R G = G_smith(NdotO, NdotI, mat.roughness);
    R denom = R(4) * NdotO * NdotI;
    V3<R> brdf = V3<R>{F.x * D * G, F.y * D * G, F.z * D * G} / denom;
    V3<R> le = scene_material<R>(s, st.material_index).emissive;
    R dsq = distance_sq(n0, nl);
    V3<R> d = ((((brdf * le) * cos_r) * cos_e) / pdf) / dsq;
    return {M<R>::fmax(d.x, R(0)), M<R>::fmax(d.y, R(0)), M<R>::fmax(d.z, R(0))};
}

// ---------------------------------------------------------------------------
// R2, R7, R9, R13, R15: primary ray, hit shading, path loop, accumulation
// ---------------------------------------------------------------------------
// path_tracer.wgsl:63-75
template <class R>
Ray<R> primary_ray(const pulse_camera &cam, uint32_t x, uint32_t y) {
    float uvx = static_cast<float>(x) / static_cast<float>(cam.width);
    float uvy = static_cast<float>(y) / static_cast<float>(cam.height);
    R cx = static_cast<R>(uvx) * R(2) - R(1);
    R cy = static_cast<R>(uvy) * R(2) - R(1);
    R h[4];
    mat_vec4<R>(cam.inverse_view_proj, cx, -cy, R(1), R(1), h);
    Ray<R> r;
    r.origin = load3<R>(cam.world_position);
    r.dir = normalize(V3<R>{h[0] / h[3], h[1] / h[3], h[2] / h[3]} - r.origin);
    r.t = sentinel<R>();
    return r;
}

template <class R>
struct HitShading {
    V3<R> normal, position;
    uint32_t material_index;
};
// path_tracer.wgsl:89-98
template <class R>
HitShading<R> shade_hit(const Scene &s, const Ray<R> &ray) {
    const pulse_mesh_instance &mi = s.instances[ray.instance_index];
    uint32_t t_idx = s.triangle_indices[mi.index_offset + ray.triangle_index];
    const pulse_triangle_data &td = s.triangle_data[mi.triangle_offset + t_idx];
    R w = R(1) - (ray.u + ray.v);
    V3<R> normal = ((w * load3<R>(td.n_first)) + (ray.u * load3<R>(td.n_second))) + (ray.v * load3<R>(td.n_third));
    HitShading<R> h;
    h.normal = normalize(transform_direction<R>(mi.object_world, normal));
    h.position = ray.origin + ray.t * ray.dir;
    h.material_index = mi.material_index;
    return h;
}

// Russian roulette, path_tracer.wgsl:104-108; returns false to stop.
template <class R>
bool russian_roulette(V3<R> &throughput, uint32_t &rng, DecisionTape &tape) {
    R p = max3(throughput);
    R e = static_cast<R>(rand_f(rng));
    if (tape.take_bool(e > p)) return false;
    if (tape.rr_mode == DecisionTape::RECORD_RR) tape.rr->push_back(static_cast<double>(p));
    if (tape.rr_mode == DecisionTape::FREEZE_RR && tape.rr_pos < tape.rr->size()) {
        const double q = (*tape.rr)[tape.rr_pos++];
        const R own = p / static_cast<R>(tape.lr);
        if (q <= 1.0) {
            tape.lr *= static_cast<double>(own) / q;
            p = static_cast<R>(q);
        } else {
            p = own;
        }
    }
    R ip = R(1) / p;
    throughput = throughput * ip;
    return true;
}

// One path (one sample of one pixel).  Modes: see pulse_hip.h.
template <class R>
V3<R> trace_path(const Scene &s, const pulse_camera &cam, const pulse_params &prm, uint32_t x, uint32_t y,
                 uint32_t accumulation_count, Counters &c, DecisionTape &tape) {
    uint32_t pixel_index = x + y * cam.width;
    uint32_t rng = pixel_seed(pixel_index, accumulation_count);
    Ray<R> ray = primary_ray<R>(cam, x, y);
    V3<R> throughput{R(1), R(1), R(1)};
    V3<R> color{R(0), R(0), R(0)};
    const bool have_lights = s.uniform.light_count > 0;
    const bool light_power = (prm.flags & PULSE_FLAG_LIGHT_POWER) != 0;

    c.primary_rays++;
    if (prm.mode == PULSE_MODE_REF_NEE_LAMBERT) {
        // path_traceBBB (path_tracer.wgsl:128-202) with the G-buffer sample
        // replaced by the camera ray's first hit.
        { CatScope scope(c, 0); trace_closest(s, ray, c, tape); }
        if (ray.t >= sentinel<R>()) return color;
        HitShading<R> h0 = shade_hit(s, ray);
        MaterialR<R> m0 = scene_material<R>(s, h0.material_index);
        R a0 = static_cast<R>(rand_f(rng));
        R a1 = static_cast<R>(rand_f(rng));
        V3<R> scatter = sample_cosine_hemisphere(h0.normal, a0, a1);
        Ray<R> next;
        next.origin = h0.position + R(0.001f) * h0.normal;
        next.dir = scatter;
        next.t = sentinel<R>();
        V3<R> direct{R(0), R(0), R(0)};
        if (have_lights) direct = sample_direct_light(s, h0.position, h0.normal, m0.base, rng, c, tape, light_power);
        color = m0.emissive + direct;
        throughput = m0.base;
        ray = next;
        for (uint32_t depth = 0; depth < prm.max_depth; ++depth) {
            { CatScope scope(c, 1); trace_closest(s, ray, c, tape); }
            if (ray.t >= sentinel<R>()) break;
            HitShading<R> h = shade_hit(s, ray);
            MaterialR<R> m = scene_material<R>(s, h.material_index);
            V3<R> dl{R(0), R(0), R(0)};
            if (have_lights) dl = sample_direct_light(s, h.position, h.normal, m.base, rng, c, tape, light_power);
            color = color + throughput * dl;
            throughput = throughput * m.base;
            if (depth >= prm.rr_start_depth && !russian_roulette(throughput, rng, tape)) break;
            R b0 = static_cast<R>(rand_f(rng));
            R b1 = static_cast<R>(rand_f(rng));
            V3<R> sd = sample_cosine_hemisphere(h.normal, b0, b1);
            ray.dir = sd;
            ray.origin = h.position + R(0.001f) * h.normal;
            ray.t = sentinel<R>();
        }
        return color;
    }

    for (uint32_t depth = 0; depth < prm.max_depth; ++depth) {
        { CatScope scope(c, depth == 0 ? 0 : 1); trace_closest(s, ray, c, tape); }
        if (ray.t >= sentinel<R>()) break;
        HitShading<R> h = shade_hit(s, ray);
        MaterialR<R> m = scene_material<R>(s, h.material_index);
        V3<R> wo = -ray.dir;
        if (prm.mode == PULSE_MODE_REF_GGX) {
            color = color + throughput * m.emissive;
        } else {
            if (depth == 0) color = color + throughput * m.emissive;
            if (have_lights) {
                V3<R> dl = prm.nee_eval == PULSE_NEE_GGX
                               ? sample_direct_light_ggx(s, h.position, h.normal, m, wo, rng, c, tape, light_power)
                               : sample_direct_light(s, h.position, h.normal, m.base, rng, c, tape, light_power);
                color = color + throughput * dl;
            }
        }
        SampleResult<R> smp = importance_sample_ggx_d(h.normal, wo, m, rng, tape);
        // optional horizon clamp (not in the reference): a path whose view or
        // sampled direction lies below the shading horizon ends here
        if ((prm.flags & PULSE_FLAG_HORIZON_CLAMP) &&
            tape.take_bool(!(smp.NdotO > R(0)) || !(smp.NdotI > R(0))))
            break;
        throughput = throughput * smp.reflectance;
        if (depth >= prm.rr_start_depth && !russian_roulette(throughput, rng, tape)) break;
        ray.dir = normalize(smp.wi);
        ray.origin = h.position + R(0.001f) * h.normal;
        ray.t = sentinel<R>();
    }
    return color;
}

// The realtime GI kernel `gi` (src/pulse/gi.wgsl:56-126) for one pixel whose
// first hit comes from the G-buffer (position, normal, base colour,
// emissive): the RNG is seeded once per pixel with pixel_index * 5817321 and
// runs on across the `spp` samples; per sample a cosine scatter from the
// G-buffer point (2 draws), Lambert NEE there (4), then up to max_depth
// bounces of: closest hit, NEE, throughput *= base colour, Russian roulette
// at every depth (1 draw), cosine scatter (2); the output is the running sum
// of colour * (1 / spp).
template <class R>
V3<R> trace_gi_pixel(const Scene &s, const pulse_params &prm, uint32_t pixel_index, V3<R> pos, V3<R> nrm,
                     V3<R> base, V3<R> emissive, Counters &c, DecisionTape &tape) {
    uint32_t rng = pixel_index * 5817321u;
    const bool have_lights = s.uniform.light_count > 0;
    const bool light_power = (prm.flags & PULSE_FLAG_LIGHT_POWER) != 0;
    const uint32_t spp = std::max(prm.spp, 1u);
    const R spp_inv = static_cast<R>(1.0f / static_cast<float>(spp));
    V3<R> out{R(0), R(0), R(0)};
    for (uint32_t sample = 0; sample < spp; ++sample) {
        R a0 = static_cast<R>(rand_f(rng));
        R a1 = static_cast<R>(rand_f(rng));
        Ray<R> ray;
        ray.dir = sample_cosine_hemisphere(nrm, a0, a1);
        ray.origin = pos + R(0.001f) * nrm;
        ray.t = sentinel<R>();
        V3<R> direct{R(0), R(0), R(0)};
        if (have_lights) direct = sample_direct_light(s, pos, nrm, base, rng, c, tape, light_power);
        V3<R> color = emissive + direct;
        V3<R> throughput = base;
        c.primary_rays++;
        for (uint32_t depth = 0; depth < prm.max_depth; ++depth) {
            { CatScope scope(c, 1); trace_closest(s, ray, c, tape); }
            if (ray.t >= sentinel<R>()) break;
            HitShading<R> h = shade_hit(s, ray);
            MaterialR<R> m = scene_material<R>(s, h.material_index);
            V3<R> dl{R(0), R(0), R(0)};
            if (have_lights) dl = sample_direct_light(s, h.position, h.normal, m.base, rng, c, tape, light_power);
            color = color + throughput * dl;
            throughput = throughput * m.base;
            if (!russian_roulette(throughput, rng, tape)) break;
            R b0 = static_cast<R>(rand_f(rng));
            R b1 = static_cast<R>(rand_f(rng));
            ray.dir = sample_cosine_hemisphere(h.normal, b0, b1);
            ray.origin = h.position + R(0.001f) * h.normal;
            ray.t = sentinel<R>();
        }
        out = out + color * spp_inv;
    }
    return out;
}

// Running mean, path_tracer.wgsl:116-123
template <class R>
inline V3<R> accumulate(V3<R> old, V3<R> color, uint32_t n) {
    if (n == 0) old = {R(0), R(0), R(0)};
    R weight = R(1) / (static_cast<R>(static_cast<float>(n)) + R(1));
    R omw = R(1) - weight;
    return {old.x * omw + color.x * weight, old.y * omw + color.y * weight, old.z * omw + color.z * weight};
}

}  // namespace oracle
