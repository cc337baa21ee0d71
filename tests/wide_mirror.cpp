// wide_mirror.cpp -- test helper (CPU): builds the quantised 4-wide BVHs with
// the library's own host code (pulse_amd/csrc/host/wide_bvh.cpp) and walks
// them the way device/wide.h does -- padded quantised slabs, near-first
// order, the per-instance closest hit and the world-t merge -- with the
// oracle's triangle test and transforms.  It checks the layout and the walk
// without a GPU; the kernels themselves are checked against the oracle by
// the GPU tests.
#include <array>
#include <cstdio>

#include "../oracle/pulse_oracle.hpp"
#include "../pulse_amd/csrc/host/wide_bvh.h"

using namespace oracle;
using pulse_host::kQuantNodeWords;
using pulse_host::kWideEmpty;
using pulse_host::WideTree;

namespace {

float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
float rcp_clamped(float v) { return 1.0f / (std::fabs(v) < 0x1p-80f ? std::copysign(0x1p-80f, v) : v); }

struct Frame {
    float sr[3], bn[3], bf[3];
    bool neg[3];
};

// As quant_frame / quant_slab (device/wide.h).
Frame frame(const uint32_t *q, V3<float> o, V3<float> r) {
    Frame f;
    const float oc[3] = {o.x, o.y, o.z}, rc[3] = {r.x, r.y, r.z};
    for (int a = 0; a < 3; ++a) {
        const float s = u2f(((q[3] >> (8 * a)) & 0xffu) << 23);
        f.sr[a] = s * rc[a];
        const float b = (u2f(q[a]) - oc[a]) * rc[a];
        const float b1 = std::fma(-1024.0f, f.sr[a], b);  // the planes enter as 1024 + q
        const float m = std::fma(2048.0f, std::fabs(f.sr[a]), std::fabs(b));
        f.bn[a] = std::fma(-0x1p-20f, m, b1);
        f.bf[a] = std::fma(0x1p-20f, m, b1);
        f.neg[a] = rc[a] < 0.0f;
    }
    return f;
}
float slab(const Frame &f, float bound, const uint32_t *q, int k) {
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        // entry plane: the low one unless the direction is negative
        const float qn = 1024.0f + static_cast<float>((q[(f.neg[a] ? 5 : 4) + 2 * a] >> (8 * k)) & 0xffu);
        const float qf = 1024.0f + static_cast<float>((q[(f.neg[a] ? 4 : 5) + 2 * a] >> (8 * k)) & 0xffu);
        lo[a] = std::fma(qn, f.sr[a], f.bn[a]);
        hi[a] = std::fma(qf, f.sr[a], f.bf[a]);
    }
    const float tmin = std::fmax(std::fmax(lo[0], lo[1]), lo[2]);
    const float tmax = std::fmin(std::fmin(hi[0], hi[1]), hi[2]);
    return (tmax >= tmin && tmin < bound && tmax > 0.0f) ? tmin : kTFar;
}

struct Layout {
    std::vector<uint32_t> tlas, blas, root;
    std::vector<uint32_t> perm;      // wide slot -> triangle slot (the refined leaves' order)
    std::vector<uint32_t> leaf_end;  // triangle slot -> end slot of its binary leaf
    uint32_t need = 0;
};

// Per triangle slot: p0, e1, e2 in f32 (the kernels' triangle rows, 12
// floats a slot), and the refinement's leaf reference counts.
std::vector<float> tri_rows(const Scene &s, size_t n_slots) {
    std::vector<float> rows(12 * std::max<size_t>(n_slots, 1), 0.0f);
    for (size_t i = 0; i < s.n_instances; ++i) {
        const pulse_mesh_instance &mi = s.instances[i];
        for (uint32_t k = 0; k < mi.triangle_count; ++k) {
            const size_t g = mi.index_offset + k;
            const pulse_primitive &p = s.primitives[mi.triangle_offset + s.triangle_indices[g]];
            float *r = &rows[12 * g];
            for (int a = 0; a < 3; ++a) {
                r[a] = p.p_first[a];
                r[3 + a] = p.p_second[a] - p.p_first[a];
                r[6 + a] = p.p_third[a] - p.p_first[a];
            }
        }
    }
    return rows;
}
size_t slot_count(const Scene &s) {
    size_t n = 0;
    for (size_t i = 0; i < s.n_instances; ++i)
        n = std::max<size_t>(n, static_cast<size_t>(s.instances[i].index_offset) + s.instances[i].triangle_count);
    return n;
}
std::vector<uint8_t> leaf_refs(const Scene &s, size_t n_slots, std::vector<uint32_t> *leaf_end = nullptr) {
    std::vector<uint8_t> refs(n_slots, 0);
    if (leaf_end) leaf_end->assign(n_slots, 0u);
    std::vector<std::pair<uint32_t, uint32_t>> seen;
    for (size_t i = 0; i < s.n_instances; ++i) {
        const uint32_t off = s.instances[i].node_offset, ioff = s.instances[i].index_offset;
        if (std::find(seen.begin(), seen.end(), std::make_pair(off, ioff)) != seen.end()) continue;
        seen.push_back({off, ioff});
        pulse_host::count_leaf_refs(reinterpret_cast<const uint32_t *>(s.blas_nodes + off), s.n_blas - off, ioff, refs,
                                    leaf_end);
    }
    return refs;
}

// The TLAS's instance boxes as the library makes them at upload: each
// instance's BLAS cut at kCutDepth, transformed.
std::vector<float> instance_boxes(const Scene &s) {
    std::vector<float> ib(6 * std::max<size_t>(s.n_instances, 1));
    for (size_t i = 0; i < s.n_instances; ++i) {
        std::vector<float> cut;
        pulse_host::blas_cut_boxes(reinterpret_cast<const uint32_t *>(s.blas_nodes + s.instances[i].node_offset),
                                   pulse_host::kCutDepth, cut);
        pulse_host::instance_world_box(s.instances[i].object_world, cut.data(), cut.size() / 6, &ib[6 * i]);
    }
    return ib;
}

bool build(const Scene &s, Layout &L) {
    WideTree t;
    const bool empty = s.tlas_nodes[0].a_or_first_instance == 0u && s.tlas_nodes[0].instance_count == 0u;
    if (empty) {
        if (!pulse_host::collapse_wide(reinterpret_cast<const uint32_t *>(s.tlas_nodes), s.n_tlas, 0u, t)) return false;
    } else {
        const std::vector<float> ib = instance_boxes(s);
        if (!pulse_host::wide_tlas(reinterpret_cast<const uint32_t *>(s.tlas_nodes), s.instance_indices, ib.data(), t))
            return false;
    }
    if (!pulse_host::quantize_wide(t, L.tlas)) return false;
    uint32_t blas_need = 0;
    std::vector<std::array<uint32_t, 3>> built;
    L.root.assign(s.n_instances, 0);
    const size_t n_slots = slot_count(s);
    const std::vector<float> rows = tri_rows(s, n_slots);
    const std::vector<uint8_t> refs = leaf_refs(s, n_slots, &L.leaf_end);
    L.perm.resize(n_slots);
    for (size_t g = 0; g < n_slots; ++g) L.perm[g] = static_cast<uint32_t>(g);
    for (size_t i = 0; i < s.n_instances; ++i) {
        const uint32_t off = s.instances[i].node_offset, ioff = s.instances[i].index_offset;
        bool found = false;
        for (auto &b : built)
            if (b[0] == off && b[1] == ioff) {
                L.root[i] = b[2];
                found = true;
            }
        if (found) continue;
        WideTree w;
        const uint32_t r = static_cast<uint32_t>(L.blas.size() / kQuantNodeWords);
        std::vector<uint32_t> nodes;
        pulse_host::refine_leaves(reinterpret_cast<const uint32_t *>(s.blas_nodes + off), s.n_blas - off, ioff,
                                  rows.data(), pulse_host::kRefineLeaf, refs, nodes, L.perm.data());
        if (!pulse_host::collapse_wide(nodes.data(), nodes.size() / 8, ioff, w) || !pulse_host::quantize_wide(w, L.blas))
            return false;
        blas_need = std::max(blas_need, w.stack_need);
        built.push_back({off, ioff, r});
        L.root[i] = r;
    }
    L.need = t.stack_need + blas_need + 1;
    return true;
}

// Near-first walk of one level: node entries from `nodes`, leaf entries to
// `leaf(first, count)`; `bound()` is the current culling bound.
template <class Bound, class Leaf>
void walk(const uint32_t *nodes, V3<float> o, V3<float> r, Bound bound, Leaf leaf, std::vector<uint32_t> &stk,
          uint64_t &visits, uint64_t &max_sp) {
    const size_t base = stk.size();
    uint32_t e = 0;
    for (;;) {
        if (e >> 31) {
            leaf(e & ((1u << 26) - 1u), ((e >> 26) & 31u) + 1u);
        } else {
            ++visits;
            const uint32_t *q = nodes + static_cast<size_t>(kQuantNodeWords) * e;
            const Frame f = frame(q, o, r);
            float d[4];
            uint32_t en[4];
            for (int k = 0; k < 4; ++k) {
                en[k] = q[10 + k];
                d[k] = en[k] != kWideEmpty ? slab(f, bound(), q, k) : kTFar;
            }
            const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
            for (auto &p : net)
                if (d[p[1]] < d[p[0]]) {
                    std::swap(d[p[0]], d[p[1]]);
                    std::swap(en[p[0]], en[p[1]]);
                }
            if (d[0] != kTFar) {
                for (int k = 3; k >= 1; --k)
                    if (d[k] != kTFar) stk.push_back(en[k]);
                max_sp = std::max<uint64_t>(max_sp, stk.size());
                e = en[0];
                continue;
            }
        }
        if (stk.size() == base) return;
        e = stk.back();
        stk.pop_back();
    }
}

void trace(const Scene &s, const Layout &L, Ray<float> &ray, uint64_t &visits, uint64_t &max_sp) {
    if (s.tlas_nodes[0].a_or_first_instance == 0u && s.tlas_nodes[0].instance_count == 0u) return;
    std::vector<uint32_t> stk;
    Counters c;
    const V3<float> wr{rcp_clamped(ray.dir.x), rcp_clamped(ray.dir.y), rcp_clamped(ray.dir.z)};
    auto enter = [&](uint32_t first, uint32_t count) {
        for (uint32_t j = first; j < first + count; ++j) {
            const uint32_t inst = s.instance_indices[j];
            const pulse_mesh_instance &mi = s.instances[inst];
            Ray<float> ro;
            ro.origin = transform_position<float>(mi.world_object, ray.origin);
            const V3<float> dd = transform_direction<float>(mi.world_object, ray.dir);
            const float dlen = length(dd);
            ro.dir = dd / dlen;
            ro.t = cull_bound(ray.origin, ray.t, dlen);  // the kernels' conservative start bound
            const float start = ro.t;
            ro.instance_index = inst;
            ro.triangle_index = 0;
            ro.u = ro.v = 0.0f;
            const V3<float> r{rcp_clamped(ro.dir.x), rcp_clamped(ro.dir.y), rcp_clamped(ro.dir.z)};
            const uint32_t *nodes = L.blas.data() + static_cast<size_t>(kQuantNodeWords) * L.root[inst];
            walk(nodes, ro.origin, r, [&] { return ro.t; },
                 [&](uint32_t f, uint32_t n) {
                     // leaf entries hold global wide slots: the mesh's slot is the record;
                     // of hits at equal t in one binary leaf the lower slot is kept (the
                     // reference tests a leaf in slot order)
                     for (uint32_t k = f; k < f + n; ++k) {
                         const uint32_t local = L.perm[k] - mi.index_offset;
                         if (local < ro.triangle_index && ro.triangle_index + mi.index_offset < L.leaf_end[L.perm[k]]) {
                             const float saved = ro.t, above = std::nextafter(saved, INFINITY);
                             ro.t = above;
                             ray_triangle_intersect(s, ro, local, inst, c);
                             if (ro.t == above) ro.t = saved;
                         } else {
                             ray_triangle_intersect(s, ro, local, inst, c);
                         }
                     }
                 },
                 stk, visits, max_sp);
            if (ro.t == start) continue;  // no hit below the bound: as a miss
            const V3<float> hw = transform_position<float>(mi.object_world, ro.origin + ro.t * ro.dir);
            const float new_t = length(hw - ray.origin);
            if (new_t < ray.t) {
                ray.t = new_t;
                ray.instance_index = inst;
                ray.triangle_index = ro.triangle_index;
                ray.u = ro.u;
                ray.v = ro.v;
            }
        }
    };
    walk(L.tlas.data(), ray.origin, wr, [&] { return ray.t; }, enter, stk, visits, max_sp);
}

Scene view(const pulse_scene_desc *d) {
    Scene s;
    s.uniform = d->uniform;
    s.primitives = static_cast<const pulse_primitive *>(d->primitives.data);
    s.triangle_data = static_cast<const pulse_triangle_data *>(d->triangle_data.data);
    s.triangle_indices = static_cast<const uint32_t *>(d->triangle_indices.data);
    s.blas_nodes = static_cast<const pulse_blas_node *>(d->blas_nodes.data);
    s.tlas_nodes = static_cast<const pulse_tlas_node *>(d->tlas_nodes.data);
    s.instance_indices = static_cast<const uint32_t *>(d->instance_indices.data);
    s.instances = static_cast<const pulse_mesh_instance *>(d->instances.data);
    s.materials = static_cast<const pulse_material *>(d->materials.data);
    s.n_primitives = d->primitives.bytes / sizeof(pulse_primitive);
    s.n_instances = d->instances.bytes / sizeof(pulse_mesh_instance);
    s.n_tlas = d->tlas_nodes.bytes / sizeof(pulse_tlas_node);
    s.n_blas = d->blas_nodes.bytes / sizeof(pulse_blas_node);
    return s;
}

}  // namespace

extern "C" {

// Closest hits of `n` rays over the 4-wide layout; stats = {node visits,
// deepest stack, the layout's stack bound}.  Returns -1 if the layout could
// not be built.
int wide_mirror_trace(const pulse_scene_desc *d, const pulse_ray *rays, pulse_hit *hits, uint32_t n,
                      uint64_t *stats) {
    const Scene s = view(d);
    Layout L;
    if (!build(s, L)) return -1;
    uint64_t visits = 0, max_sp = 0;
    for (uint32_t i = 0; i < n; ++i) {
        Ray<float> r;
        r.origin = load3<float>(rays[i].origin);
        r.dir = load3<float>(rays[i].dir);
        r.t = rays[i].t_max;
        r.instance_index = r.triangle_index = 0;
        r.u = r.v = 0.0f;
        trace(s, L, r, visits, max_sp);
        hits[i] = pulse_hit{r.t, r.instance_index, r.triangle_index, r.u, r.v};
    }
    stats[0] = visits;
    stats[1] = max_sp;
    stats[2] = L.need;
    return 0;
}

// Layout invariants over the whole scene: every binary leaf's primitive
// range appears, in order and exactly once, as wide leaf entries of at most
// 32; every quantised child box contains its exact box (decoded in double).
// out = {leaf ranges checked, containment failures, coverage failures}.
int wide_mirror_check(const pulse_scene_desc *d, uint64_t *out) {
    const Scene s = view(d);
    out[0] = out[1] = out[2] = 0;
    const size_t n_slots = slot_count(s);
    const std::vector<float> rows = tri_rows(s, n_slots);
    const std::vector<uint8_t> refs = leaf_refs(s, n_slots);
    std::vector<uint32_t> perm(n_slots);
    for (size_t g = 0; g < n_slots; ++g) perm[g] = static_cast<uint32_t>(g);
    auto check_tree = [&](const uint32_t *bin, size_t nb, uint32_t off, bool tlas) {
        WideTree t;
        std::vector<uint32_t> q, refined;
        const bool empty_tlas = tlas && bin[3] == 0 && bin[7] == 0;
        if (!tlas) pulse_host::refine_leaves(bin, nb, off, rows.data(), pulse_host::kRefineLeaf, refs, refined, perm.data());
        if (tlas && !empty_tlas) {
            // the library's TLAS: a surface-area tree over the same instances
            const std::vector<float> ib = instance_boxes(s);
            if (!pulse_host::wide_tlas(bin, s.instance_indices, ib.data(), t)) return false;
        } else if (!pulse_host::collapse_wide(tlas ? bin : refined.data(), tlas ? nb : refined.size() / 8, off, t)) {
            return false;
        }
        if (tlas && !empty_tlas) {
            const std::vector<float> ib = instance_boxes(s);
            for (size_t i = 0; i < s.n_instances; ++i) {
                const float *b = &ib[6 * i];
                // every triangle vertex of the instance, transformed, lies in its box
                const pulse_mesh_instance &mi = s.instances[i];
                for (uint32_t k = 0; k < mi.triangle_count; k += 7) {
                    const pulse_primitive &pr = s.primitives[mi.triangle_offset + k];
                    for (const float *v : {pr.p_first, pr.p_second, pr.p_third}) {
                        const V3<float> w = transform_position<float>(mi.object_world, load3<float>(v));
                        if (w.x < b[0] || w.y < b[1] || w.z < b[2] || w.x > b[3] || w.y > b[4] || w.z > b[5]) ++out[1];
                    }
                }
            }
        }
        if (!pulse_host::quantize_wide(t, q)) return false;
        const size_t nodes = t.words.size() / pulse_host::kWideNodeWords;
        {
            // the refit's level order: every node once, every child in an earlier level
            std::vector<uint32_t> lv;
            pulse_host::wide_levels(t, lv);
            const uint32_t n_levels = lv[0];
            std::vector<int> level_of(nodes, -1);
            for (uint32_t l = 0; l < n_levels; ++l)
                for (uint32_t j = lv[1 + l]; j < lv[2 + l]; ++j) {
                    const uint32_t i = lv[2 + n_levels + j];
                    if (i >= nodes || level_of[i] != -1) ++out[2];
                    else level_of[i] = static_cast<int>(l);
                }
            for (size_t i = 0; i < nodes; ++i) {
                if (level_of[i] < 0) {
                    ++out[2];
                    continue;
                }
                for (int k = 0; k < 4; ++k) {
                    const uint32_t e = t.words[i * pulse_host::kWideNodeWords + 24 + k];
                    if (e != kWideEmpty && !(e >> 31) && (e >= nodes || level_of[e] >= level_of[i])) ++out[2];
                }
            }
        }
        std::vector<std::pair<uint32_t, uint32_t>> wide_leaves;
        for (size_t i = 0; i < nodes; ++i) {
            const uint32_t *w = t.words.data() + i * pulse_host::kWideNodeWords;
            const uint32_t *qq = q.data() + i * kQuantNodeWords;
            for (int k = 0; k < 4; ++k) {
                const uint32_t e = w[24 + k];
                if (e == kWideEmpty) continue;
                if (qq[10 + k] != e) ++out[1];
                if (e >> 31) {
                    const uint32_t first = e & ((1u << 26) - 1u), count = ((e >> 26) & 31u) + 1u;
                    wide_leaves.push_back({first, count});
                    // a BLAS leaf's f32 box holds its triangles (p0, p0 + e1, p0 + e2 in
                    // double; the reference's own leaf boxes hold p1 and p2, which lie
                    // within 2^-24 |e| of those)
                    for (uint32_t j = first; j < first + count && !tlas; ++j) {
                        const float *r = &rows[12 * static_cast<size_t>(perm[j])];
                        for (int a = 0; a < 3; ++a) {
                            float lo, hi;
                            std::memcpy(&lo, w + 8 * a + k, 4);
                            std::memcpy(&hi, w + 8 * a + 4 + k, 4);
                            const double v0 = r[a], v1 = v0 + static_cast<double>(r[3 + a]), v2 = v0 + static_cast<double>(r[6 + a]);
                            const double tol = std::ldexp(std::fabs(r[3 + a]) + std::fabs(r[6 + a]), -23);
                            if (std::min(v0, std::min(v1, v2)) < lo - tol || std::max(v0, std::max(v1, v2)) > hi + tol)
                                ++out[1];
                        }
                    }
                }
                for (int a = 0; a < 3; ++a) {
                    float lo, hi;
                    std::memcpy(&lo, w + 8 * a + k, 4);
                    std::memcpy(&hi, w + 8 * a + 4 + k, 4);
                    const double sc = std::ldexp(1.0, static_cast<int>((qq[3] >> (8 * a)) & 0xffu) - 127);
                    const double org = u2f(qq[a]);
                    const double ql = (qq[4 + 2 * a] >> (8 * k)) & 0xffu, qh = (qq[5 + 2 * a] >> (8 * k)) & 0xffu;
                    if (org + ql * sc > lo || org + qh * sc < hi) ++out[1];
                }
            }
        }
        // binary leaves in node order must be covered by the wide leaf chunks
        std::vector<std::pair<uint32_t, uint32_t>> bin_leaves;
        for (size_t i = 0; i < nb; ++i) {
            const uint32_t *w = bin + 8 * i;
            if (w[7] > 0) bin_leaves.push_back({w[3] + off, w[7]});
        }
        if (nb == 1 && bin[7] == 0 && bin[3] == 0) bin_leaves.clear();  // empty tree
        std::vector<uint32_t> covered_b, covered_w;
        for (auto &l : bin_leaves) {
            for (uint32_t j = 0; j < l.second; ++j) covered_b.push_back(l.first + j);
            // the refinement reorders slots within the leaf only
            if (!tlas) {
                std::vector<uint32_t> moved(perm.begin() + l.first, perm.begin() + l.first + l.second);
                std::sort(moved.begin(), moved.end());
                for (uint32_t j = 0; j < l.second; ++j)
                    if (moved[j] != l.first + j) ++out[2];
            }
        }
        for (auto &l : wide_leaves) {
            if (l.second > 32) ++out[2];
            for (uint32_t j = 0; j < l.second; ++j) covered_w.push_back(l.first + j);
        }
        std::sort(covered_b.begin(), covered_b.end());
        std::sort(covered_w.begin(), covered_w.end());
        if (covered_b != covered_w) ++out[2];
        out[0] += wide_leaves.size();
        return true;
    };
    if (!check_tree(reinterpret_cast<const uint32_t *>(s.tlas_nodes), s.n_tlas, 0u, true)) return -1;
    std::vector<std::pair<uint32_t, uint32_t>> seen;
    for (size_t i = 0; i < s.n_instances; ++i) {
        const uint32_t off = s.instances[i].node_offset, ioff = s.instances[i].index_offset;
        if (std::find(seen.begin(), seen.end(), std::make_pair(off, ioff)) != seen.end()) continue;
        seen.push_back({off, ioff});
        // the mesh's BLAS spans the nodes up to the next mesh's (or the end)
        size_t end = s.n_blas;
        for (size_t j = 0; j < s.n_instances; ++j)
            if (s.instances[j].node_offset > off) end = std::min<size_t>(end, s.instances[j].node_offset);
        if (!check_tree(reinterpret_cast<const uint32_t *>(s.blas_nodes + off), end - off, ioff, false)) return -1;
    }
    return 0;
}

}  // extern "C"
