// wide_bvh.cpp -- collapse of the reference's binary BVHs into 4-wide nodes
// (layout in wide_bvh.h).  Which binary interior nodes become wide nodes is
// chosen by dynamic programming over the binary tree for the least summed
// surface area of the wide nodes (the expected number of node visits of a
// ray through the root box, up to a constant): f(n, k), the least cost of
// covering the subtree of n by at most k subtrees, is 0 for a leaf and
// min(f(n, 1), min_i f(a, i) + f(b, k - i)) otherwise, with f(n, 1) =
// area(n) + min_i f(a, i) + f(b, 4 - i) (n itself a wide node over its
// cheapest cover by four).  The leaves, their order and their primitive
// ranges are the binary tree's.

#include "wide_bvh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <thread>

namespace pulse_host {
namespace {

struct Item {
    float lo[3], hi[3];
    bool interior;   // binary interior node `bin` (else a primitive range)
    uint32_t bin;
    uint32_t first, count;
};

struct Builder {
    const uint32_t *nodes;
    size_t n;
    uint32_t first_offset;
    WideTree &out;
    bool ok = true;
    // per binary node: cover cost f(n, k) for k = 1..4 (index k - 1) and
    // the left child's budget of the best split for k = 2..4 (0: n is
    // its own cover, k >= 2) and for n's own wide node (index 0)
    std::vector<std::array<float, 4>> cost;
    std::vector<std::array<uint8_t, 4>> split;

    Item read(uint32_t i) const {
        Item it{};
        const uint32_t *w = nodes + 8 * static_cast<size_t>(i);
        std::memcpy(it.lo, w, 12);
        std::memcpy(it.hi, w + 4, 12);
        it.bin = i;
        it.first = w[3];
        it.count = w[7];
        it.interior = it.count == 0;
        return it;
    }
    static float half_area(const Item &it) {
        float dx = it.hi[0] - it.lo[0], dy = it.hi[1] - it.lo[1], dz = it.hi[2] - it.lo[2];
        return dx * dy + dy * dz + dz * dx;
    }

    uint32_t alloc() {
        uint32_t idx = static_cast<uint32_t>(out.words.size() / kWideNodeWords);
        out.words.resize(out.words.size() + kWideNodeWords, 0u);
        for (int k = 0; k < 4; ++k) out.words[static_cast<size_t>(idx) * kWideNodeWords + 24 + k] = kWideEmpty;
        return idx;
    }
    void set_child(uint32_t node, int k, const Item &it, uint32_t entry) {
        uint32_t *w = out.words.data() + static_cast<size_t>(node) * kWideNodeWords;
        for (int a = 0; a < 3; ++a) {
            std::memcpy(w + 8 * a + k, &it.lo[a], 4);
            std::memcpy(w + 8 * a + 4 + k, &it.hi[a], 4);
        }
        w[24 + k] = entry;
    }
    uint32_t leaf_entry(uint32_t first, uint32_t count) {
        const uint64_t f = static_cast<uint64_t>(first) + first_offset;
        if (f >= kWideFirstLimit) ok = false;
        return kWideLeafBit | ((count - 1u) << 26) | (static_cast<uint32_t>(f) & ((1u << 26) - 1u));
    }

    // Children that become wide nodes get their nodes side by side, right
    // after each other, before any of their own children: siblings share
    // 128-byte lines and every node still precedes its children.
    static constexpr uint32_t kNoNode = ~0u;
    static bool needs_node(const Item &it) { return it.interior || it.count > kWideMaxLeaf; }

    // Node over a primitive range of more than 32 (chunks of 32 in order),
    // at `node` (kNoNode: a new one); returns its index, `need` = the stack
    // entries below it.
    uint32_t range_node(const Item &box, uint32_t first, uint32_t count, uint32_t &need, uint32_t node = kNoNode) {
        if (node == kNoNode) node = alloc();
        uint32_t chunks = (count + kWideMaxLeaf - 1) / kWideMaxLeaf;
        if (chunks <= 4) {
            for (uint32_t c = 0; c < chunks; ++c) {
                uint32_t f = first + c * kWideMaxLeaf;
                set_child(node, static_cast<int>(c), box, leaf_entry(f, std::min(kWideMaxLeaf, first + count - f)));
            }
            need = chunks - 1;
            return node;
        }
        // four groups of whole chunks
        uint32_t per = (chunks + 3) / 4, deepest = 0;
        uint32_t pre[4] = {kNoNode, kNoNode, kNoNode, kNoNode};
        int k = 0;
        for (uint32_t c = 0; c < chunks; c += per, ++k)
            if (std::min(per * kWideMaxLeaf, first + count - (first + c * kWideMaxLeaf)) > kWideMaxLeaf) pre[k] = alloc();
        k = 0;
        for (uint32_t c = 0; c < chunks; c += per, ++k) {
            uint32_t f = first + c * kWideMaxLeaf;
            uint32_t cnt = std::min(per * kWideMaxLeaf, first + count - f);
            uint32_t sub_need = 0, entry;
            if (cnt <= kWideMaxLeaf) {
                entry = leaf_entry(f, cnt);
            } else {
                entry = range_node(box, f, cnt, sub_need, pre[k]);
            }
            set_child(node, k, box, entry);
            deepest = std::max(deepest, sub_need);
        }
        need = static_cast<uint32_t>(k - 1) + deepest;
        return node;
    }

    // Entry for a child item (its wide node at `node` if it needs one,
    // kNoNode: a new one); `need` = stack entries below it.
    uint32_t entry_of(const Item &it, uint32_t &need, uint32_t node = kNoNode) {
        need = 0;
        if (!it.interior) {
            if (it.count <= kWideMaxLeaf) return leaf_entry(it.first, it.count);
            return range_node(it, it.first, it.count, need, node);
        }
        return interior_node(it.bin, need, node);
    }

    // The cover tables for the subtree of `root` (post order, iterative).
    void plan(uint32_t root) {
        cost.assign(n, std::array<float, 4>{0.0f, 0.0f, 0.0f, 0.0f});
        split.assign(n, std::array<uint8_t, 4>{0, 0, 0, 0});
        std::vector<std::pair<uint32_t, bool>> st{{root, false}};
        while (!st.empty()) {
            auto [i, done] = st.back();
            st.pop_back();
            const Item it = read(i);
            if (!it.interior || it.first + 1u >= n || it.first == 0) continue;  // leaf (or a malformed entry: no cost)
            if (!done) {
                st.push_back({i, true});
                st.push_back({it.first + 1u, false});
                st.push_back({it.first, false});
                continue;
            }
            const auto &fa = cost[it.first], &fb = cost[it.first + 1u];
            auto best_split = [&](int k, uint8_t &arg) {
                float best = 0.0f;
                arg = 0;
                for (int i2 = 1; i2 < k; ++i2) {
                    const float c = fa[i2 - 1] + fb[k - i2 - 1];
                    if (arg == 0 || c < best) {
                        best = c;
                        arg = static_cast<uint8_t>(i2);
                    }
                }
                return best;
            };
            auto &f = cost[i];
            auto &sp = split[i];
            f[0] = half_area(it) + best_split(4, sp[0]);
            for (int k = 2; k <= 4; ++k) {
                uint8_t arg;
                const float g = best_split(k, arg);
                if (g < f[0]) {
                    f[k - 1] = g;
                    sp[k - 1] = arg;
                } else {
                    f[k - 1] = f[0];
                    sp[k - 1] = 0;
                }
            }
        }
    }
    // Appends the cheapest cover of the subtree of `i` by at most k items.
    void cover(uint32_t i, int k, std::vector<Item> &items) {
        const Item it = read(i);
        if (!it.interior || k == 1 || split[i][k - 1] == 0) {
            items.push_back(it);
            return;
        }
        const int a = split[i][k - 1];
        cover(it.first, a, items);
        cover(it.first + 1u, k - a, items);
    }

    // Wide node for binary interior node `b`, at `node` (kNoNode: a new one).
    uint32_t interior_node(uint32_t b, uint32_t &need, uint32_t node = kNoNode) {
        Item root = read(b);
        std::vector<Item> items;
        const int a = split[b][0] ? split[b][0] : 2;  // 0: not planned (checked inputs always are)
        cover(root.first, a, items);
        cover(root.first + 1u, 4 - a, items);
        if (node == kNoNode) node = alloc();
        uint32_t pre[4] = {kNoNode, kNoNode, kNoNode, kNoNode};
        for (size_t k = 0; k < items.size(); ++k)
            if (needs_node(items[k])) pre[k] = alloc();
        uint32_t deepest = 0;
        for (size_t k = 0; k < items.size(); ++k) {
            uint32_t sub = 0;
            uint32_t e = entry_of(items[k], sub, pre[k]);
            set_child(node, static_cast<int>(k), items[k], e);
            deepest = std::max(deepest, sub);
        }
        need = static_cast<uint32_t>(items.size() - 1) + deepest;
        return node;
    }
};

}  // namespace

bool collapse_wide(const uint32_t *nodes, size_t n, uint32_t first_offset, WideTree &out) {
    out.words.clear();
    out.stack_need = 0;
    Builder b{nodes, n, first_offset, out, true, {}, {}};
    if (n == 0) {
        b.alloc();
        return true;
    }
    Item root = b.read(0);
    if (root.interior && root.first == 0) {  // empty tree {0, 0}: one node without children
        b.alloc();
        return true;
    }
    uint32_t need = 0;
    if (root.interior) {
        b.plan(0);
        b.interior_node(0, need);
    } else {
        uint32_t node = b.alloc();
        uint32_t sub = 0;
        uint32_t e = b.entry_of(root, sub);
        b.set_child(node, 0, root, e);
        need = sub;
    }
    out.stack_need = need;
    // every interior entry names a later node: a walk visits each node at
    // most once per entry into the tree (the traversal has no visit guard)
    const size_t nodes_out = out.words.size() / kWideNodeWords;
    for (size_t i = 0; i < nodes_out && b.ok; ++i)
        for (int k = 0; k < 4; ++k) {
            const uint32_t e = out.words[i * kWideNodeWords + 24 + k];
            if (e != kWideEmpty && !(e & kWideLeafBit) && (e <= i || e >= nodes_out)) b.ok = false;
        }
    return b.ok;
}

void blas_cut_boxes(const uint32_t *nodes, int depth, std::vector<float> &out) {
    std::vector<std::pair<uint32_t, int>> st{{0u, 0}};
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        const uint32_t *w = nodes + 8 * static_cast<size_t>(i);
        if (w[7] == 0 && d < depth && w[3] != 0) {
            st.push_back({w[3] + 1u, d + 1});
            st.push_back({w[3], d + 1});
            continue;
        }
        float b[6];
        std::memcpy(b, w, 12);
        std::memcpy(b + 3, w + 4, 12);
        out.insert(out.end(), b, b + 6);
    }
}

void instance_world_box(const float *m, const float *boxes, size_t n, float out[6]) {
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (size_t j = 0; j < n; ++j) {
        const float *b = boxes + 6 * j;
        for (int c = 0; c < 8; ++c) {
            const float px = (c & 1) ? b[3] : b[0], py = (c & 2) ? b[4] : b[1], pz = (c & 4) ? b[5] : b[2];
            const float hx = m[0] * px + m[4] * py + m[8] * pz + m[12] * 1.0f;
            const float hy = m[1] * px + m[5] * py + m[9] * pz + m[13] * 1.0f;
            const float hz = m[2] * px + m[6] * py + m[10] * pz + m[14] * 1.0f;
            const float hw = m[3] * px + m[7] * py + m[11] * pz + m[15] * 1.0f;
            const float w[3] = {hx / hw, hy / hw, hz / hw};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], w[a]);
                hi[a] = std::max(hi[a], w[a]);
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        out[a] = lo[a] - std::fabs(lo[a]) * 0x1p-20f;
        out[3 + a] = hi[a] + std::fabs(hi[a]) * 0x1p-20f;
    }
}

void instance_world_boxes(const float *m, size_t m_stride, const float *cut_boxes, const uint32_t *cut_range,
                          size_t n, float *out, int threads) {
    auto run = [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i)
            instance_world_box(m + m_stride * i, cut_boxes + 6 * static_cast<size_t>(cut_range[2 * i]),
                               cut_range[2 * i + 1], out + 6 * i);
    };
    const size_t t = std::max<size_t>(1, std::min<size_t>(static_cast<size_t>(std::max(threads, 1)), n / 64));
    if (t <= 1) {
        run(0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t k = 0; k < t; ++k) pool.emplace_back(run, n * k / t, n * (k + 1) / t);
    for (auto &th : pool) th.join();
}

namespace {

struct SahBox {
    float lo[3], hi[3];
    void grow(const float *b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b[a]);
            hi[a] = std::max(hi[a], b[3 + a]);
        }
    }
    double half_area() const {
        const double dx = static_cast<double>(hi[0]) - lo[0], dy = static_cast<double>(hi[1]) - lo[1],
                     dz = static_cast<double>(hi[2]) - lo[2];
        return dx * dy + dy * dz + dz * dx;
    }
};
constexpr SahBox kEmptyBox = {{3.402823466e38f, 3.402823466e38f, 3.402823466e38f},
                              {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f}};
constexpr size_t kSweepItems = 4096;  // above: 64 centroid bins per axis
constexpr int kSahBins = 64;

}  // namespace

void sah_tree(const float *boxes, size_t n, std::vector<uint32_t> &nodes) {
    nodes.clear();
    if (n == 0) return;
    std::vector<uint32_t> items(n);
    std::vector<float> cen(3 * n);
    for (size_t j = 0; j < n; ++j) {
        items[j] = static_cast<uint32_t>(j);
        for (int a = 0; a < 3; ++a) cen[3 * j + a] = 0.5f * boxes[6 * j + a] + 0.5f * boxes[6 * j + 3 + a];
    }
    auto put = [&](uint32_t node, const SahBox &b, uint32_t first, uint32_t count) {
        uint32_t *w = nodes.data() + 8 * static_cast<size_t>(node);
        std::memcpy(w, b.lo, 12);
        std::memcpy(w + 4, b.hi, 12);
        w[3] = first;
        w[7] = count;
    };
    struct Task {
        uint32_t node, begin, end;
    };
    nodes.assign(8, 0u);
    std::vector<Task> st{{0u, 0u, static_cast<uint32_t>(n)}};
    std::vector<double> right_area(n + 1);
    while (!st.empty()) {
        const Task t = st.back();
        st.pop_back();
        const uint32_t cnt = t.end - t.begin;
        SahBox box = kEmptyBox;
        for (uint32_t i = t.begin; i < t.end; ++i) box.grow(boxes + 6 * static_cast<size_t>(items[i]));
        if (cnt == 1) {
            put(t.node, box, items[t.begin], 1u);
            continue;
        }
        // best split: left = the first k items in centroid order along an axis
        int best_axis = -1;
        uint32_t best_k = 0;
        double best_cost = 0.0;
        float best_plane = 0.0f;
        if (cnt <= kSweepItems) {
            std::vector<uint32_t> order(items.begin() + t.begin, items.begin() + t.end);
            for (int a = 0; a < 3; ++a) {
                std::stable_sort(order.begin(), order.end(),
                                 [&](uint32_t x, uint32_t y) { return cen[3 * x + a] < cen[3 * y + a]; });
                SahBox r = kEmptyBox;
                for (uint32_t k = cnt; k-- > 1;) {
                    r.grow(boxes + 6 * static_cast<size_t>(order[k]));
                    right_area[k] = r.half_area() * (cnt - k);
                }
                SahBox l = kEmptyBox;
                for (uint32_t k = 1; k < cnt; ++k) {
                    l.grow(boxes + 6 * static_cast<size_t>(order[k - 1]));
                    const double c = l.half_area() * k + right_area[k];
                    if (best_axis < 0 || c < best_cost) {
                        best_axis = a;
                        best_k = k;
                        best_cost = c;
                    }
                }
            }
            std::stable_sort(items.begin() + t.begin, items.begin() + t.end,
                             [&](uint32_t x, uint32_t y) { return cen[3 * x + best_axis] < cen[3 * y + best_axis]; });
        } else {
            SahBox cb = kEmptyBox;
            for (uint32_t i = t.begin; i < t.end; ++i) {
                const float *c = &cen[3 * static_cast<size_t>(items[i])];
                const float b6[6] = {c[0], c[1], c[2], c[0], c[1], c[2]};
                cb.grow(b6);
            }
            for (int a = 0; a < 3; ++a) {
                const float ext = cb.hi[a] - cb.lo[a];
                if (!(ext > 0.0f)) continue;
                SahBox bb[kSahBins];
                uint32_t bc[kSahBins] = {};
                for (int k = 0; k < kSahBins; ++k) bb[k] = kEmptyBox;
                auto bin_of = [&](uint32_t it) {
                    const int k = static_cast<int>((cen[3 * it + a] - cb.lo[a]) / ext * kSahBins);
                    return std::min(std::max(k, 0), kSahBins - 1);
                };
                for (uint32_t i = t.begin; i < t.end; ++i) {
                    const int k = bin_of(items[i]);
                    bb[k].grow(boxes + 6 * static_cast<size_t>(items[i]));
                    bc[k] += 1;
                }
                double ra[kSahBins];
                uint32_t rc[kSahBins];
                SahBox r = kEmptyBox;
                uint32_t rn = 0;
                for (int k = kSahBins - 1; k > 0; --k) {
                    if (bc[k]) {
                        const float *lo = bb[k].lo, *hi = bb[k].hi;
                        const float b6[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
                        r.grow(b6);
                    }
                    rn += bc[k];
                    ra[k] = rn ? r.half_area() : 0.0;
                    rc[k] = rn;
                }
                SahBox l = kEmptyBox;
                uint32_t ln = 0;
                for (int k = 1; k < kSahBins; ++k) {
                    if (bc[k - 1]) {
                        const float *lo = bb[k - 1].lo, *hi = bb[k - 1].hi;
                        const float b6[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
                        l.grow(b6);
                    }
                    ln += bc[k - 1];
                    if (ln == 0 || rc[k] == 0) continue;
                    const double c = l.half_area() * ln + ra[k] * rc[k];
                    if (best_axis < 0 || c < best_cost) {
                        best_axis = a;
                        best_k = ln;
                        best_cost = c;
                        best_plane = cb.lo[a] + ext * static_cast<float>(k) / kSahBins;
                    }
                }
            }
            if (best_axis >= 0) {
                const int a = best_axis;
                auto mid = std::stable_partition(items.begin() + t.begin, items.begin() + t.end, [&](uint32_t it) {
                    const float ext = cb.hi[a] - cb.lo[a];
                    const int k = std::min(std::max(static_cast<int>((cen[3 * it + a] - cb.lo[a]) / ext * kSahBins), 0),
                                           kSahBins - 1);
                    return cb.lo[a] + ext * static_cast<float>(k) / kSahBins < best_plane;
                });
                best_k = static_cast<uint32_t>(mid - (items.begin() + t.begin));
            }
        }
        if (best_axis < 0 || best_k == 0 || best_k >= cnt) best_k = cnt / 2;  // coincident centroids: halves
        const uint32_t c0 = static_cast<uint32_t>(nodes.size() / 8);
        nodes.resize(nodes.size() + 16, 0u);
        put(t.node, box, c0, 0u);
        st.push_back({c0 + 1u, t.begin + best_k, t.end});
        st.push_back({c0, t.begin, t.begin + best_k});
    }
}

bool wide_tlas(const uint32_t *tlas, const uint32_t *iidx, const float *inst_box, WideTree &out) {
    std::vector<uint32_t> pos, st{0u};
    while (!st.empty()) {
        const uint32_t *w = tlas + 8 * static_cast<size_t>(st.back());
        st.pop_back();
        if (w[7] == 0) {
            st.push_back(w[3] + 1u);
            st.push_back(w[3]);
        } else {
            for (uint32_t c = 0; c < w[7]; ++c) pos.push_back(w[3] + c);
        }
    }
    std::vector<float> pb(6 * pos.size());
    for (size_t j = 0; j < pos.size(); ++j) std::memcpy(&pb[6 * j], inst_box + 6 * static_cast<size_t>(iidx[pos[j]]), 24);
    std::vector<uint32_t> nodes;
    sah_tree(pb.data(), pos.size(), nodes);
    for (size_t i = 0; i < nodes.size() / 8; ++i)
        if (nodes[8 * i + 7] != 0) nodes[8 * i + 3] = pos[nodes[8 * i + 3]];
    return collapse_wide(nodes.data(), nodes.size() / 8, 0u, out);
}

void wide_levels(const WideTree &t, std::vector<uint32_t> &out) {
    const size_t n = t.words.size() / kWideNodeWords;
    std::vector<uint32_t> depth(n, 0u), frontier{0u};
    uint32_t deepest = 0;
    while (!frontier.empty() && n) {
        std::vector<uint32_t> next;
        for (uint32_t i : frontier)
            for (int k = 0; k < 4; ++k) {
                const uint32_t e = t.words[static_cast<size_t>(i) * kWideNodeWords + 24 + k];
                if (e == kWideEmpty || (e & kWideLeafBit) || e >= n) continue;
                depth[e] = depth[i] + 1u;
                deepest = std::max(deepest, depth[e]);
                next.push_back(e);
            }
        frontier.swap(next);
    }
    const uint32_t levels = n ? deepest + 1u : 0u;
    out.assign(2u + levels + n, 0u);
    out[0] = levels;
    std::vector<uint32_t> count(levels, 0u);
    for (size_t i = 0; i < n; ++i) count[deepest - depth[i]] += 1u;  // level 0: the deepest
    for (uint32_t l = 0; l < levels; ++l) out[2 + l] = out[1 + l] + count[l];
    std::vector<uint32_t> fill(out.begin() + 1, out.begin() + 1 + levels);
    for (size_t i = 0; i < n; ++i) out[2 + levels + fill[deepest - depth[i]]++] = static_cast<uint32_t>(i);
}

bool quantize_wide(const WideTree &in, std::vector<uint32_t> &out) {
    const size_t n = in.words.size() / kWideNodeWords;
    const size_t base = out.size();
    out.resize(base + n * kQuantNodeWords, 0u);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t *w = in.words.data() + i * kWideNodeWords;
        uint32_t *q = out.data() + base + i * kQuantNodeWords;
        bool valid[4];
        for (int k = 0; k < 4; ++k) {
            valid[k] = w[24 + k] != kWideEmpty;
            q[10 + k] = w[24 + k];
        }
        uint32_t exps = 0;
        for (int a = 0; a < 3; ++a) {
            float clo[4], chi[4];
            std::memcpy(clo, w + 8 * a, 16);
            std::memcpy(chi, w + 8 * a + 4, 16);
            float lo = 0.0f, hi = 0.0f;
            bool any = false;
            for (int k = 0; k < 4; ++k) {
                if (!valid[k]) continue;
                if (!std::isfinite(clo[k]) || !std::isfinite(chi[k]) || std::fabs(clo[k]) > 0x1p40f ||
                    std::fabs(chi[k]) > 0x1p40f)
                    return false;
                lo = any ? std::min(lo, clo[k]) : clo[k];
                hi = any ? std::max(hi, chi[k]) : chi[k];
                any = true;
            }
            const double extent = static_cast<double>(hi) - lo;
            int e = -126;
            if (extent > 0.0) {
                e = std::max(-126, static_cast<int>(std::ceil(std::log2(extent / 255.0))));
                while (extent > 255.0 * std::ldexp(1.0, e)) ++e;
            }
            if (e > 127) return false;
            const double sc = std::ldexp(1.0, e);
            exps |= static_cast<uint32_t>(e + 127) << (8 * a);
            std::memcpy(q + a, &lo, 4);
            uint32_t qlo = 0, qhi = 0;
            for (int k = 0; k < 4; ++k) {
                if (!valid[k]) continue;
                double l = std::floor((static_cast<double>(clo[k]) - lo) / sc);
                double h = std::ceil((static_cast<double>(chi[k]) - lo) / sc);
                while (l > 0.0 && lo + l * sc > clo[k]) l -= 1.0;
                while (h < 255.0 && lo + h * sc < chi[k]) h += 1.0;
                if (l < 0.0 || h > 255.0 || lo + l * sc > clo[k] || lo + h * sc < chi[k]) return false;
                qlo |= static_cast<uint32_t>(l) << (8 * k);
                qhi |= static_cast<uint32_t>(h) << (8 * k);
            }
            q[4 + 2 * a] = qlo;
            q[5 + 2 * a] = qhi;
        }
        q[3] = exps;
    }
    return true;
}

namespace {

struct TriBox {
    float lo[3], hi[3];
    double c[3];  // centroid
    uint32_t slot;
};

float round_down(double x) {
    float f = static_cast<float>(x);
    return static_cast<double>(f) > x ? std::nextafter(f, -INFINITY) : f;
}
float round_up(double x) {
    float f = static_cast<float>(x);
    return static_cast<double>(f) < x ? std::nextafter(f, INFINITY) : f;
}

TriBox tri_box(const float *tri, uint32_t slot) {
    const float *t = tri + 12 * static_cast<size_t>(slot);
    TriBox b{};
    b.slot = slot;
    double lo[3], hi[3], ext = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double v0 = t[a], v1 = v0 + static_cast<double>(t[3 + a]), v2 = v0 + static_cast<double>(t[6 + a]);
        lo[a] = std::min(v0, std::min(v1, v2));
        hi[a] = std::max(v0, std::max(v1, v2));
        b.c[a] = (v0 + v1 + v2) / 3.0;
        ext = std::max(ext, hi[a] - lo[a]);
    }
    for (int a = 0; a < 3; ++a) {
        const double pad = std::ldexp(ext, -8) + std::ldexp(std::max(std::fabs(lo[a]), std::fabs(hi[a])), -20);
        b.lo[a] = round_down(lo[a] - pad);
        b.hi[a] = round_up(hi[a] + pad);
    }
    return b;
}

double half_area(const float lo[3], const float hi[3]) {
    const double dx = static_cast<double>(hi[0]) - lo[0], dy = static_cast<double>(hi[1]) - lo[1],
                 dz = static_cast<double>(hi[2]) - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// Fills node `at` of `out` over items[0..cnt), whose triangles take the
// slots first_offset + pos .. + cnt in the order of the leaves.
void build_sub(std::vector<uint32_t> &out, uint32_t at, TriBox *items, uint32_t cnt, uint32_t pos,
               uint32_t max_leaf, uint32_t first_offset, uint32_t *perm) {
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = items[0].lo[a];
        hi[a] = items[0].hi[a];
        for (uint32_t k = 1; k < cnt; ++k) {
            lo[a] = std::min(lo[a], items[k].lo[a]);
            hi[a] = std::max(hi[a], items[k].hi[a]);
        }
    }
    {
        uint32_t *w = out.data() + 8 * static_cast<size_t>(at);
        std::memcpy(w, lo, 12);
        std::memcpy(w + 4, hi, 12);
    }
    if (cnt <= max_leaf) {
        out[8 * static_cast<size_t>(at) + 3] = pos;
        out[8 * static_cast<size_t>(at) + 7] = cnt;
        for (uint32_t k = 0; k < cnt; ++k) perm[static_cast<size_t>(first_offset) + pos + k] = items[k].slot;
        return;
    }
    // least area-weighted cost over the centroid orders of the three axes
    int best_axis = 0;
    uint32_t best_k = cnt / 2;
    double best = INFINITY;
    std::vector<double> right(cnt);
    for (int a = 0; a < 3; ++a) {
        std::sort(items, items + cnt, [a](const TriBox &x, const TriBox &y) {
            return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.slot < y.slot);
        });
        float rl[3], rh[3];
        for (uint32_t k = cnt; k-- > 1;) {
            for (int b = 0; b < 3; ++b) {
                rl[b] = k == cnt - 1 ? items[k].lo[b] : std::min(rl[b], items[k].lo[b]);
                rh[b] = k == cnt - 1 ? items[k].hi[b] : std::max(rh[b], items[k].hi[b]);
            }
            right[k] = half_area(rl, rh) * (cnt - k);
        }
        float ll[3], lh[3];
        for (uint32_t k = 1; k < cnt; ++k) {
            for (int b = 0; b < 3; ++b) {
                ll[b] = k == 1 ? items[0].lo[b] : std::min(ll[b], items[k - 1].lo[b]);
                lh[b] = k == 1 ? items[0].hi[b] : std::max(lh[b], items[k - 1].hi[b]);
            }
            const double cost = half_area(ll, lh) * k + right[k];
            if (cost < best) {
                best = cost;
                best_axis = a;
                best_k = k;
            }
        }
    }
    std::sort(items, items + cnt, [best_axis](const TriBox &x, const TriBox &y) {
        return x.c[best_axis] < y.c[best_axis] || (x.c[best_axis] == y.c[best_axis] && x.slot < y.slot);
    });
    const uint32_t child = static_cast<uint32_t>(out.size() / 8);
    out.resize(out.size() + 16, 0u);
    out[8 * static_cast<size_t>(at) + 3] = child;
    out[8 * static_cast<size_t>(at) + 7] = 0u;
    build_sub(out, child, items, best_k, pos, max_leaf, first_offset, perm);
    build_sub(out, child + 1, items + best_k, cnt - best_k, pos + best_k, max_leaf, first_offset, perm);
}

}  // namespace

void count_leaf_refs(const uint32_t *nodes, size_t n, uint32_t first_offset, std::vector<uint8_t> &refs,
                     std::vector<uint32_t> *leaf_end) {
    if (n == 0) return;
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        const uint32_t *w = nodes + 8 * static_cast<size_t>(i);
        if (w[7] == 0) {
            if (w[3] != 0 && static_cast<size_t>(w[3]) + 1 < n) {
                st.push_back(w[3]);
                st.push_back(w[3] + 1u);
            }
            continue;
        }
        const size_t end = static_cast<size_t>(first_offset) + w[3] + w[7];
        for (uint32_t k = 0; k < w[7]; ++k) {
            const size_t g = static_cast<size_t>(first_offset) + w[3] + k;
            if (g < refs.size() && refs[g] < 255) ++refs[g];
            if (leaf_end && g < leaf_end->size()) (*leaf_end)[g] = static_cast<uint32_t>(end);
        }
    }
}

void refine_leaves(const uint32_t *nodes, size_t n, uint32_t first_offset, const float *tri, uint32_t max_leaf,
                   const std::vector<uint8_t> &refs, std::vector<uint32_t> &out, uint32_t *perm) {
    out.clear();
    if (n == 0) return;
    size_t top = 0;
    std::vector<uint32_t> leaves, st{0u};
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        top = std::max<size_t>(top, i);
        const uint32_t *w = nodes + 8 * static_cast<size_t>(i);
        if (w[7] == 0) {
            if (w[3] != 0 && static_cast<size_t>(w[3]) + 1 < n) {
                st.push_back(w[3]);
                st.push_back(w[3] + 1u);
            }
        } else {
            leaves.push_back(i);
        }
    }
    out.assign(nodes, nodes + 8 * (top + 1));
    std::vector<TriBox> items;
    for (uint32_t leaf : leaves) {
        const uint32_t first = nodes[8 * static_cast<size_t>(leaf) + 3], count = nodes[8 * static_cast<size_t>(leaf) + 7];
        if (count <= max_leaf || count > kRefineMaxLeaf) continue;
        const size_t g0 = static_cast<size_t>(first_offset) + first;
        if (g0 + count > refs.size()) continue;
        bool own = true;
        for (uint32_t k = 0; k < count && own; ++k) own = refs[g0 + k] == 1;
        if (!own) continue;
        items.clear();
        for (uint32_t k = 0; k < count; ++k) items.push_back(tri_box(tri, static_cast<uint32_t>(g0 + k)));
        build_sub(out, leaf, items.data(), count, first, max_leaf, first_offset, perm);
    }
}

}  // namespace pulse_host
