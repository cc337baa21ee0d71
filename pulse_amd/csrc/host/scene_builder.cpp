// Host-side producer of the scene buffer layout: binned-SAH BLAS/TLAS
// builders and the scene preparation that concatenates meshes, computes
// instance bounds and the light tables.
//
// Restates src/scene/blas.rs, src/scene/tlas.rs and src/scene/mod.rs:372-742
// in C++ with the same f32 arithmetic (compiled with -ffp-contract=off, so
// every product and sum rounds exactly as the Rust code does) and the same
// node emission order, so the node arrays are byte-identical for the same
// input.  Quirks kept on purpose:
//   * bins and the running SAH boxes start from AABB::default(), i.e. the
//     origin is inside every box (blas.rs:21-25, 164-167, 213-214);
//   * no SAH leaf-cost test; leaves at <= 8 triangles / <= 1 instance
//     (blas.rs:96, tlas.rs:73);
//   * an axis whose centroid range is empty is skipped; if every axis is
//     empty the split is (axis 0, position 0.0) (blas.rs:176-190).
// One deviation: when the root node's partition sends every item right the
// reference's u32 `j -= 1` underflows and panics (blas.rs:120, tlas.rs:411);
// here that node simply stays a leaf, which is what every other node does in
// the same situation (blas.rs:124-128).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "common.h"

namespace {

struct V3 {
    float x, y, z;
};

inline float comp(const V3 &v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
// f32::min / f32::max (glam Vec3::min / max are component-wise f32::min / max)
inline V3 vmin(V3 a, V3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 vscale(float s, V3 v) { return {s * v.x, s * v.y, s * v.z}; }
inline float vlength(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline float vlength_sq(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

constexpr float kFMax = std::numeric_limits<float>::max();

// AABB of blas.rs:21-42 / tlas.rs:452-473; Default is the degenerate box at
// the origin.
struct Aabb {
    V3 mn{0.f, 0.f, 0.f};
    V3 mx{0.f, 0.f, 0.f};
    void grow(V3 p) {
        mn = vmin(mn, p);
        mx = vmax(mx, p);
    }
    void grow(const Aabb &b) {
        grow(b.mn);
        grow(b.mx);
    }
    float area() const {
        V3 e = vsub(mx, mn);
        return e.x * e.y + e.y * e.z + e.z * e.x;
    }
};

// Rust `f32 as usize`: truncation toward zero, saturating, NaN -> 0.
inline size_t sat_usize(float f) {
    if (!(f > 0.f)) return 0;
    if (f >= 1.8446744e19f) return std::numeric_limits<size_t>::max();
    return static_cast<size_t>(f);
}

struct BuildNode {
    V3 mn, mx;
    uint32_t a_or_first;
    uint32_t count;
};

// Item adaptor for the shared builder: triangles (BLAS) or instances (TLAS).
struct TriItems {
    const pulse_primitive *prims;
    std::vector<V3> centroids;
    static constexpr uint32_t kLeafMax = 8;   // blas.rs:96
    static constexpr int kBins = 20;          // blas.rs:193
    V3 p(uint32_t tri, int k) const {
        const float *q = k == 0 ? prims[tri].p_first : (k == 1 ? prims[tri].p_second : prims[tri].p_third);
        return {q[0], q[1], q[2]};
    }
    V3 centroid(uint32_t item) const { return centroids[item]; }
    // calculate_node_aabb, blas.rs:296-314
    void node_bounds(const std::vector<uint32_t> &idx, BuildNode &n) const {
        n.mn = {kFMax, kFMax, kFMax};
        n.mx = {-kFMax, -kFMax, -kFMax};
        for (uint32_t i = 0; i < n.count; ++i) {
            uint32_t t = idx[n.a_or_first + i];
            for (int k = 0; k < 3; ++k) {
                n.mn = vmin(n.mn, p(t, k));
            }
            for (int k = 0; k < 3; ++k) {
                n.mx = vmax(n.mx, p(t, k));
            }
        }
    }
    // bins[..].bounds.grow_position(triangle.positions[0..3]), blas.rs:203-205
    void grow_bin(Aabb &b, uint32_t item) const {
        for (int k = 0; k < 3; ++k) b.grow(p(item, k));
    }
};

struct InstItems {
    const pulse_instance_bounds *inst;
    static constexpr uint32_t kLeafMax = 1;   // tlas.rs:73
    static constexpr int kBins = 2;           // tlas.rs:192
    V3 centroid(uint32_t item) const {
        const float *c = inst[item].center;
        return {c[0], c[1], c[2]};
    }
    V3 bmin(uint32_t item) const { return {inst[item].bounds_min[0], inst[item].bounds_min[1], inst[item].bounds_min[2]}; }
    V3 bmax(uint32_t item) const { return {inst[item].bounds_max[0], inst[item].bounds_max[1], inst[item].bounds_max[2]}; }
    // calculate_node_aabb, tlas.rs:251-263
    void node_bounds(const std::vector<uint32_t> &idx, BuildNode &n) const {
        n.mn = {kFMax, kFMax, kFMax};
        n.mx = {-kFMax, -kFMax, -kFMax};
        for (uint32_t i = n.a_or_first; i < n.a_or_first + n.count; ++i) {
            n.mn = vmin(n.mn, bmin(idx[i]));
            n.mx = vmax(n.mx, bmax(idx[i]));
        }
    }
    // tlas.rs:200-201
    void grow_bin(Aabb &b, uint32_t item) const {
        b.grow(bmin(item));
        b.grow(bmax(item));
    }
};

struct Split {
    int axis;
    float position;
    float cost;
};

// find_best_split_plane, blas.rs:170-253 / tlas.rs:168-249.
template <class Items>
Split find_best_split_plane(const Items &items, const BuildNode &node, const std::vector<uint32_t> &idx) {
    constexpr int BC = Items::kBins;
    Split best{0, 0.0f, 1e30f};
    for (int axis = 0; axis < 3; ++axis) {
        float bounds_min = 1e30f;
        float bounds_max = -1e30f;
        for (uint32_t i = 0; i < node.count; ++i) {
            float c = comp(items.centroid(idx[node.a_or_first + i]), axis);
            bounds_min = std::fmin(bounds_min, c);
            bounds_max = std::fmax(bounds_max, c);
        }
        if (bounds_min == bounds_max) continue;

        Aabb bins[BC];
        uint32_t bin_count[BC] = {};
        const float bin_size_inv = static_cast<float>(BC) / (bounds_max - bounds_min);
        for (uint32_t i = 0; i < node.count; ++i) {
            uint32_t item = idx[node.a_or_first + i];
            float c = comp(items.centroid(item), axis);
            size_t b = std::min(static_cast<size_t>(BC - 1), sat_usize((c - bounds_min) * bin_size_inv));
            bin_count[b] += 1;
            items.grow_bin(bins[b], item);
        }

        float area_a[BC - 1], area_b[BC - 1];
        uint32_t count_a[BC - 1], count_b[BC - 1];
        Aabb box_a, box_b;
        uint32_t sum_a = 0, sum_b = 0;
        for (int i = 0; i < BC - 1; ++i) {
            sum_a += bin_count[i];
            count_a[i] = sum_a;
            box_a.grow(bins[i]);
            area_a[i] = box_a.area();

            sum_b += bin_count[BC - 1 - i];
            count_b[BC - 2 - i] = sum_b;
            box_b.grow(bins[BC - 1 - i]);
            area_b[BC - 2 - i] = box_b.area();
        }

        const float bin_size = (bounds_max - bounds_min) / static_cast<float>(BC);
        for (int i = 0; i < BC - 1; ++i) {
            float plane_cost = static_cast<float>(count_a[i]) * area_a[i] +
                               static_cast<float>(count_b[i]) * area_b[i];
            if (plane_cost < best.cost) {
                best.axis = axis;
                best.position = bounds_min + bin_size * static_cast<float>(i + 1);
                best.cost = plane_cost;
            }
        }
    }
    return best;
}

// build_blas + subdivide (blas.rs:44-161) and build_tlas + subdivide
// (tlas.rs:29-136).  The recursion "split node, append child pair, recurse
// into a, then into b" is run with an explicit stack (b pushed below a), which
// emits nodes in the same order without recursion-depth limits.
template <class Items>
void build_tree(const Items &items, uint32_t n, std::vector<BuildNode> &nodes, std::vector<uint32_t> &idx) {
    idx.resize(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    nodes.clear();
    nodes.reserve(n > 0 ? 2 * static_cast<size_t>(n) - 1 : 1);
    BuildNode root{};
    root.a_or_first = 0;
    root.count = n;
    items.node_bounds(idx, root);
    nodes.push_back(root);

    std::vector<uint32_t> stack;
    stack.push_back(0);
    while (!stack.empty()) {
        uint32_t node_idx = stack.back();
        stack.pop_back();
        BuildNode node = nodes[node_idx];
        if (node.count <= Items::kLeafMax) continue;

        Split s = find_best_split_plane(items, node, idx);

        int64_t i = node.a_or_first;
        int64_t j = i + static_cast<int64_t>(node.count) - 1;
        while (i <= j) {
            if (comp(items.centroid(idx[static_cast<size_t>(i)]), s.axis) < s.position) {
                i += 1;
            } else {
                std::swap(idx[static_cast<size_t>(i)], idx[static_cast<size_t>(j)]);
                j -= 1;
            }
        }

        uint32_t a_count = static_cast<uint32_t>(i) - node.a_or_first;
        if (a_count == 0 || a_count == node.count) continue;

        BuildNode child_a{};
        child_a.a_or_first = node.a_or_first;
        child_a.count = a_count;
        items.node_bounds(idx, child_a);
        uint32_t child_a_index = static_cast<uint32_t>(nodes.size());
        nodes.push_back(child_a);

        BuildNode child_b{};
        child_b.a_or_first = static_cast<uint32_t>(i);
        child_b.count = node.count - a_count;
        items.node_bounds(idx, child_b);
        nodes.push_back(child_b);

        nodes[node_idx].a_or_first = child_a_index;
        nodes[node_idx].count = 0;

        stack.push_back(child_a_index + 1);
        stack.push_back(child_a_index);
    }
}

template <class NodeT>
void emit_node(const BuildNode &b, NodeT &out) {
    out.aabb_min[0] = b.mn.x;
    out.aabb_min[1] = b.mn.y;
    out.aabb_min[2] = b.mn.z;
    out.aabb_max[0] = b.mx.x;
    out.aabb_max[1] = b.mx.y;
    out.aabb_max[2] = b.mx.z;
}

// Depth (edges root -> deepest leaf) of a node array where count == 0 marks
// an interior node whose children are first and first+1.
template <class NodeT, class CountFn, class FirstFn>
uint32_t tree_depth(const NodeT *nodes, size_t n_nodes, CountFn count, FirstFn first) {
    if (n_nodes == 0) return 0;
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.push_back({0u, 0u});
    while (!st.empty()) {
        auto [ni, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        if (count(nodes[ni]) == 0) {
            uint32_t a = first(nodes[ni]);
            if (a + 1 < n_nodes && a != 0) {
                st.push_back({a, d + 1});
                st.push_back({a + 1, d + 1});
            }
        }
    }
    return best;
}

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// glam Mat4::mul_vec4 order: ((c0*x + c1*y) + c2*z) + c3*w, then xyz / w
// (transform_position, src/utilities.rs:42-45).
inline V3 transform_position(const float *m, V3 p) {
    float hx = m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * 1.0f;
    float hy = m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * 1.0f;
    float hz = m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * 1.0f;
    float hw = m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15] * 1.0f;
    return {hx / hw, hy / hw, hz / hw};
}

// General 4x4 inverse (cofactor expansion in double, rounded to f32).  The
// reference uses glam's Mat4::inverse (scene/mod.rs:596); its f32 rounding is
// not reproducible here, and the matrix is an input of the kernels, so only
// its accuracy matters.
bool invert4(const float *mf, float *outf) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = mf[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) outf[i] = static_cast<float>(inv[i] / det);
    return true;
}

// create_triangle_area_cdf, scene/mod.rs:700-724 (Heron's formula in f32).
float triangle_area_cdf(const pulse_primitive *prims, uint32_t n, std::vector<float> &cdf_out) {
    std::vector<float> areas(n);
    for (uint32_t i = 0; i < n; ++i) {
        V3 p0{prims[i].p_first[0], prims[i].p_first[1], prims[i].p_first[2]};
        V3 p1{prims[i].p_second[0], prims[i].p_second[1], prims[i].p_second[2]};
        V3 p2{prims[i].p_third[0], prims[i].p_third[1], prims[i].p_third[2]};
        float a = vlength(vsub(p1, p0));
        float b = vlength(vsub(p2, p0));
        float c = vlength(vsub(p2, p1));
        float s = 0.5f * (a + b + c);
        areas[i] = std::sqrt(s * (s - a) * (s - b) * (s - c));
    }
    float total = 0.0f;
    for (float a : areas) total += a;
    float prev = 0.0f;
    for (float a : areas) {
        float current = prev + a / total;
        cdf_out.push_back(current);
        prev = current;
    }
    return total;
}

}  // namespace

// ---------------------------------------------------------------------------
// Single-tree entry points
// ---------------------------------------------------------------------------

extern "C" int pulse_build_blas(const pulse_primitive *prims, uint32_t n, pulse_blas_node *nodes_out,
                                uint32_t nodes_capacity, uint32_t *node_count, uint32_t *tri_indices_out) {
    if (!prims || !nodes_out || !node_count || !tri_indices_out)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_build_blas: null argument");
    if (n == 0) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_build_blas: a mesh needs at least one triangle");
    if (nodes_capacity < 2 * n - 1)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_build_blas: nodes_capacity %u < 2n-1 = %u", nodes_capacity, 2 * n - 1);
    TriItems items;
    items.prims = prims;
    items.centroids.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        // blas.rs:47-63: centroid is the AABB centre of the triangle.
        V3 bmin{kFMax, kFMax, kFMax}, bmax{-kFMax, -kFMax, -kFMax};
        for (int k = 0; k < 3; ++k) bmin = vmin(bmin, items.p(i, k));
        for (int k = 0; k < 3; ++k) bmax = vmax(bmax, items.p(i, k));
        items.centroids[i] = vadd(bmin, vscale(0.5f, vsub(bmax, bmin)));
    }
    std::vector<BuildNode> nodes;
    std::vector<uint32_t> idx;
    build_tree(items, n, nodes, idx);
    for (size_t k = 0; k < nodes.size(); ++k) {
        emit_node(nodes[k], nodes_out[k]);
        nodes_out[k].a_or_first_tri = nodes[k].a_or_first;
        nodes_out[k].tri_count = nodes[k].count;
    }
    std::memcpy(tri_indices_out, idx.data(), sizeof(uint32_t) * n);
    *node_count = static_cast<uint32_t>(nodes.size());
    return PULSE_OK;
}

extern "C" int pulse_build_tlas(const pulse_instance_bounds *inst, uint32_t n, pulse_tlas_node *nodes_out,
                                uint32_t nodes_capacity, uint32_t *node_count,
                                uint32_t *instance_indices_out) {
    if (!nodes_out || !node_count || (n > 0 && (!inst || !instance_indices_out)))
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_build_tlas: null argument");
    uint32_t need = n > 0 ? 2 * n - 1 : 1;
    if (nodes_capacity < need)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_build_tlas: nodes_capacity %u < %u", nodes_capacity, need);
    if (n == 0) {
        // PulseTLASNode::invalid(), tlas.rs:13-20
        std::memset(&nodes_out[0], 0, sizeof(pulse_tlas_node));
        *node_count = 1;
        return PULSE_OK;
    }
    InstItems items{inst};
    std::vector<BuildNode> nodes;
    std::vector<uint32_t> idx;
    build_tree(items, n, nodes, idx);
    for (size_t k = 0; k < nodes.size(); ++k) {
        emit_node(nodes[k], nodes_out[k]);
        nodes_out[k].a_or_first_instance = nodes[k].a_or_first;
        nodes_out[k].instance_count = nodes[k].count;
    }
    std::memcpy(instance_indices_out, idx.data(), sizeof(uint32_t) * n);
    *node_count = static_cast<uint32_t>(nodes.size());
    return PULSE_OK;
}

// ---------------------------------------------------------------------------
// Scene builder
// ---------------------------------------------------------------------------

struct pulse_scene_builder {
    struct Mesh {
        std::vector<pulse_primitive> prims;
        std::vector<pulse_triangle_data> tdata;
        std::vector<pulse_blas_node> nodes;
        std::vector<uint32_t> tri_indices;
        uint32_t depth = 0;
        std::vector<float> area_cdf;  // light triangle-area CDF, once the mesh is used as a light
        float area_total = 0.f;
        bool has_area_cdf = false;
    };
    struct Instance {
        uint32_t mesh, material;
        float m[16];
    };
    std::vector<Mesh> meshes;
    std::vector<pulse_material> materials;
    std::vector<Instance> instances;

    // outputs
    pulse_scene_uniform uniform{};
    std::vector<pulse_primitive> primitives;
    std::vector<pulse_triangle_data> triangle_data;
    std::vector<uint32_t> triangle_indices;
    std::vector<pulse_blas_node> blas_nodes;
    std::vector<pulse_tlas_node> tlas_nodes;
    std::vector<uint32_t> instance_indices;
    std::vector<pulse_mesh_instance> out_instances;
    std::vector<float> emission_strength_cdf;
    std::vector<float> triangle_cdfs;
    std::vector<float> light_mesh_areas;
    std::vector<pulse_light_data_index> light_indices;
    pulse_scene_build_info info{};
    bool built = false;
    // where each mesh landed in the concatenated buffers (set by the build)
    struct MeshIndex {
        uint32_t triangle_offset, triangle_count, index_offset, node_offset;
    };
    std::vector<MeshIndex> mesh_index;
};

namespace {
// prepare_mesh_instances (scene/mod.rs:561-697), create_emission_strength_cdf
// (726-742) and build_tlas (693): everything that depends on the instances,
// given the meshes' places in the concatenated buffers.
int build_instances_and_tlas(pulse_scene_builder &S);
}  // namespace

extern "C" int pulse_sb_create(pulse_scene_builder **out) {
    if (!out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_create: null out");
    *out = new pulse_scene_builder();
    return PULSE_OK;
}

extern "C" int pulse_sb_destroy(pulse_scene_builder *sb) {
    delete sb;
    return PULSE_OK;
}

// prepare_extracted_mesh_assets, scene/mod.rs:372-448: one PulsePrimitive and
// one PulseTriangleData per index triple, in index-buffer order.
extern "C" int pulse_sb_add_mesh(pulse_scene_builder *sb, const float *positions, const float *normals,
                                 const float *uvs, uint32_t vertex_count, const uint32_t *indices,
                                 uint32_t index_count, uint32_t *mesh_id) {
    if (!sb || !positions || !normals || !indices)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_mesh: null argument");
    if (index_count == 0 || index_count % 3 != 0)
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_mesh: index_count %u is not a positive multiple of 3", index_count);
    pulse_scene_builder::Mesh m;
    uint32_t nt = index_count / 3;
    m.prims.resize(nt);
    m.tdata.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        pulse_primitive &p = m.prims[t];
        pulse_triangle_data &d = m.tdata[t];
        std::memset(&p, 0, sizeof(p));
        std::memset(&d, 0, sizeof(d));
        float *pp[3] = {p.p_first, p.p_second, p.p_third};
        float *nn[3] = {d.n_first, d.n_second, d.n_third};
        float *uu[3] = {d.uv_first, d.uv_second, d.uv_third};
        for (int k = 0; k < 3; ++k) {
            uint32_t v = indices[3 * t + k];
            if (v >= vertex_count)
                PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_mesh: index %u >= vertex_count %u", v, vertex_count);
            for (int c = 0; c < 3; ++c) {
                pp[k][c] = positions[3 * v + c];
                nn[k][c] = normals[3 * v + c];
            }
            uu[k][0] = uvs ? uvs[2 * v] : 0.f;
            uu[k][1] = uvs ? uvs[2 * v + 1] : 0.f;
        }
    }
    sb->meshes.push_back(std::move(m));
    if (mesh_id) *mesh_id = static_cast<uint32_t>(sb->meshes.size() - 1);
    sb->built = false;
    return PULSE_OK;
}

extern "C" int pulse_sb_add_material(pulse_scene_builder *sb, const pulse_material *m, uint32_t *material_id) {
    if (!sb || !m) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_material: null argument");
    pulse_material c = *m;
    c._pad = 0.f;
    sb->materials.push_back(c);
    if (material_id) *material_id = static_cast<uint32_t>(sb->materials.size() - 1);
    sb->built = false;
    return PULSE_OK;
}

extern "C" int pulse_sb_add_instance(pulse_scene_builder *sb, uint32_t mesh_id, uint32_t material_id,
                                     const float *transform16) {
    if (!sb || !transform16) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_instance: null argument");
    if (mesh_id >= sb->meshes.size()) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_instance: bad mesh id %u", mesh_id);
    if (material_id >= sb->materials.size())
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_add_instance: bad material id %u", material_id);
    pulse_scene_builder::Instance in;
    in.mesh = mesh_id;
    in.material = material_id;
    std::memcpy(in.m, transform16, sizeof(in.m));
    sb->instances.push_back(in);
    sb->built = false;
    return PULSE_OK;
}

extern "C" int pulse_sb_build(pulse_scene_builder *sb, int threads) {
    if (!sb) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_build: null builder");
    pulse_scene_builder &S = *sb;
    const size_t nm = S.meshes.size();

    // --- BLAS per mesh (only the node/index arrays depend on the build) ---
    double t0 = now_ms();
    std::atomic<size_t> next{0};
    std::atomic<int> failed{0};
    auto worker = [&]() {
        for (;;) {
            size_t k = next.fetch_add(1);
            if (k >= nm) break;
            auto &m = S.meshes[k];
            if (!m.nodes.empty()) continue;  // already built
            uint32_t n = static_cast<uint32_t>(m.prims.size());
            m.nodes.resize(2 * static_cast<size_t>(n) - 1);
            m.tri_indices.resize(n);
            uint32_t cnt = 0;
            if (pulse_build_blas(m.prims.data(), n, m.nodes.data(), static_cast<uint32_t>(m.nodes.size()), &cnt,
                                 m.tri_indices.data()) != PULSE_OK) {
                failed = 1;
                continue;
            }
            m.nodes.resize(cnt);
            m.depth = tree_depth(m.nodes.data(), m.nodes.size(), [](const pulse_blas_node &x) { return x.tri_count; },
                                 [](const pulse_blas_node &x) { return x.a_or_first_tri; });
        }
    };
    int nt = std::max(1, std::min<int>(threads, static_cast<int>(nm)));
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    if (failed) PULSE_FAIL(PULSE_E_BUILD, "pulse_sb_build: BLAS build failed");
    double t1 = now_ms();

    // --- prepare_mesh_data, scene/mod.rs:471-510 ---
    S.primitives.clear();
    S.triangle_data.clear();
    S.triangle_indices.clear();
    S.blas_nodes.clear();
    std::vector<pulse_scene_builder::MeshIndex> &mesh_index = S.mesh_index;
    mesh_index.assign(nm, {});
    S.info = pulse_scene_build_info{};
    for (size_t k = 0; k < nm; ++k) {
        auto &m = S.meshes[k];
        mesh_index[k] = {static_cast<uint32_t>(S.primitives.size()), static_cast<uint32_t>(m.prims.size()),
                         static_cast<uint32_t>(S.triangle_indices.size()), static_cast<uint32_t>(S.blas_nodes.size())};
        S.primitives.insert(S.primitives.end(), m.prims.begin(), m.prims.end());
        S.triangle_data.insert(S.triangle_data.end(), m.tdata.begin(), m.tdata.end());
        S.triangle_indices.insert(S.triangle_indices.end(), m.tri_indices.begin(), m.tri_indices.end());
        S.blas_nodes.insert(S.blas_nodes.end(), m.nodes.begin(), m.nodes.end());
        S.info.unique_triangles += m.prims.size();
        S.info.max_blas_depth = std::max(S.info.max_blas_depth, m.depth);
    }

    int rc = build_instances_and_tlas(S);
    if (rc != PULSE_OK) return rc;
    double t2 = now_ms();
    S.info.ms_blas = t1 - t0;
    S.info.ms_instances_tlas = t2 - t1;
    S.built = true;
    return PULSE_OK;
}

namespace {
int build_instances_and_tlas(pulse_scene_builder &S) {
    using MeshIndex = pulse_scene_builder::MeshIndex;
    const auto &mesh_index = S.mesh_index;
    const size_t nm = S.meshes.size();
    S.info.instanced_triangles = 0;
    // --- prepare_mesh_instances, scene/mod.rs:561-697 ---
    S.out_instances.clear();
    S.triangle_cdfs.clear();
    S.light_mesh_areas.clear();
    S.light_indices.clear();
    std::vector<float> strengths;
    std::vector<pulse_instance_bounds> inst_bounds;
    for (const auto &in : S.instances) {
        pulse_mesh_instance mi;
        std::memset(&mi, 0, sizeof(mi));
        std::memcpy(mi.object_world, in.m, sizeof(in.m));
        if (!invert4(in.m, mi.world_object))
            PULSE_FAIL(PULSE_E_BUILD, "pulse_sb_build: instance transform is singular");
        const MeshIndex &mx = mesh_index[in.mesh];
        mi.triangle_offset = mx.triangle_offset;
        mi.triangle_count = mx.triangle_count;
        mi.index_offset = mx.index_offset;
        mi.node_offset = mx.node_offset;
        mi.material_index = in.material;
        S.out_instances.push_back(mi);
        S.info.instanced_triangles += mx.triangle_count;

        const pulse_material &mat = S.materials[in.material];
        V3 em{mat.emissive[0], mat.emissive[1], mat.emissive[2]};
        if (vlength(em) > 0.0001f) {
            strengths.push_back(vlength_sq(em));
            // the area CDF depends on the mesh only: computed once per mesh
            auto &mesh = S.meshes[in.mesh];
            if (!mesh.has_area_cdf) {
                mesh.area_total = triangle_area_cdf(&S.primitives[mx.triangle_offset], mx.triangle_count, mesh.area_cdf);
                mesh.has_area_cdf = true;
            }
            const std::vector<float> &cdf = mesh.area_cdf;
            const float total = mesh.area_total;
            pulse_light_data_index li;
            li.cdf_offset = static_cast<uint32_t>(S.triangle_cdfs.size());
            li.mesh_instance_index = static_cast<uint32_t>(S.out_instances.size() - 1);
            S.light_indices.push_back(li);
            S.triangle_cdfs.insert(S.triangle_cdfs.end(), cdf.begin(), cdf.end());
            S.light_mesh_areas.push_back(total);
        }

        // World-space bounds from the 8 transformed corners of the BLAS root
        // (scene/mod.rs:624-680).
        const pulse_blas_node &root = S.blas_nodes[mx.node_offset];
        V3 b0{root.aabb_min[0], root.aabb_min[1], root.aabb_min[2]};
        V3 b1{root.aabb_max[0], root.aabb_max[1], root.aabb_max[2]};
        V3 corners[8] = {{b0.x, b0.y, b0.z}, {b1.x, b0.y, b0.z}, {b0.x, b1.y, b0.z}, {b0.x, b0.y, b1.z},
                         {b1.x, b1.y, b0.z}, {b0.x, b1.y, b1.z}, {b1.x, b0.y, b1.z}, {b1.x, b1.y, b1.z}};
        V3 wmin{kFMax, kFMax, kFMax}, wmax{-kFMax, -kFMax, -kFMax};
        for (auto &c : corners) wmin = vmin(wmin, transform_position(in.m, c));
        for (auto &c : corners) wmax = vmax(wmax, transform_position(in.m, c));
        V3 center = vadd(wmin, vscale(0.5f, vsub(wmax, wmin)));
        pulse_instance_bounds ib;
        ib.bounds_min[0] = wmin.x; ib.bounds_min[1] = wmin.y; ib.bounds_min[2] = wmin.z;
        ib.bounds_max[0] = wmax.x; ib.bounds_max[1] = wmax.y; ib.bounds_max[2] = wmax.z;
        ib.center[0] = center.x; ib.center[1] = center.y; ib.center[2] = center.z;
        inst_bounds.push_back(ib);
    }

    // create_emission_strength_cdf, scene/mod.rs:726-742
    S.emission_strength_cdf.clear();
    {
        float total = 0.f;
        for (float s : strengths) total += s;
        float prev = 0.f;
        for (float s : strengths) {
            float cur = prev + s / total;
            S.emission_strength_cdf.push_back(cur);
            prev = cur;
        }
    }

    // build_tlas, scene/mod.rs:693
    uint32_t ni = static_cast<uint32_t>(inst_bounds.size());
    S.tlas_nodes.assign(ni > 0 ? 2 * static_cast<size_t>(ni) - 1 : 1, pulse_tlas_node{});
    S.instance_indices.assign(ni, 0);
    uint32_t tcount = 0;
    int rc = pulse_build_tlas(inst_bounds.data(), ni, S.tlas_nodes.data(), static_cast<uint32_t>(S.tlas_nodes.size()),
                              &tcount, S.instance_indices.data());
    if (rc != PULSE_OK) return rc;
    S.tlas_nodes.resize(tcount);
    S.uniform.instance_count = ni;
    S.uniform.light_count = static_cast<uint32_t>(S.light_indices.size());
    S.info.mesh_count = static_cast<uint32_t>(nm);
    S.info.instance_count = ni;
    S.info.light_count = S.uniform.light_count;
    S.info.material_count = static_cast<uint32_t>(S.materials.size());
    S.info.blas_node_count = static_cast<uint32_t>(S.blas_nodes.size());
    S.info.tlas_node_count = tcount;
    S.info.tlas_depth = ni > 0 ? tree_depth(S.tlas_nodes.data(), S.tlas_nodes.size(),
                                            [](const pulse_tlas_node &x) { return x.instance_count; },
                                            [](const pulse_tlas_node &x) { return x.a_or_first_instance; })
                               : 0;
    return PULSE_OK;
}
}  // namespace

// Moves instance `instance_id` (column-major object->world matrix); takes
// effect at the next pulse_sb_update_instances or pulse_sb_build.
extern "C" int pulse_sb_set_instance_transform(pulse_scene_builder *sb, uint32_t instance_id, const float *transform16) {
    if (!sb || !transform16) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_set_instance_transform: null argument");
    if (instance_id >= sb->instances.size())
        PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_set_instance_transform: bad instance id %u", instance_id);
    std::memcpy(sb->instances[instance_id].m, transform16, sizeof(float) * 16);
    return PULSE_OK;
}

// Recomputes the instance records, light tables and TLAS of a built scene
// (the reference's per-frame prepare_mesh_instances + build_tlas) without
// touching the meshes, their BLAS or the concatenated triangle buffers.
extern "C" int pulse_sb_update_instances(pulse_scene_builder *sb) {
    if (!sb) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_update_instances: null builder");
    if (!sb->built) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_update_instances: call pulse_sb_build first");
    double t0 = now_ms();
    int rc = build_instances_and_tlas(*sb);
    if (rc != PULSE_OK) {
        sb->built = false;
        return rc;
    }
    sb->info.ms_instances_tlas = now_ms() - t0;
    return PULSE_OK;
}

extern "C" int pulse_sb_get_desc(pulse_scene_builder *sb, pulse_scene_desc *out) {
    if (!sb || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_get_desc: null argument");
    if (!sb->built) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_get_desc: call pulse_sb_build first");
    auto view = [](const auto &v) {
        pulse_buffer_view b;
        b.data = v.empty() ? nullptr : static_cast<const void *>(v.data());
        b.bytes = v.size() * sizeof(v[0]);
        return b;
    };
    out->uniform = sb->uniform;
    out->primitives = view(sb->primitives);
    out->triangle_data = view(sb->triangle_data);
    out->triangle_indices = view(sb->triangle_indices);
    out->blas_nodes = view(sb->blas_nodes);
    out->tlas_nodes = view(sb->tlas_nodes);
    out->instance_indices = view(sb->instance_indices);
    out->instances = view(sb->out_instances);
    out->materials = view(sb->materials);
    out->light_emission_strength_cdf = view(sb->emission_strength_cdf);
    out->light_triangle_area_cdfs = view(sb->triangle_cdfs);
    out->light_mesh_areas = view(sb->light_mesh_areas);
    out->light_indices = view(sb->light_indices);
    return PULSE_OK;
}

extern "C" int pulse_sb_get_info(pulse_scene_builder *sb, pulse_scene_build_info *out) {
    if (!sb || !out) PULSE_FAIL(PULSE_E_INVALID_ARG, "pulse_sb_get_info: null argument");
    *out = sb->info;
    return PULSE_OK;
}
