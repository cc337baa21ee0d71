// wide_bvh.h -- 4-wide BVHs derived at upload from the reference's binary
// TLAS and BLASes (same leaves, same triangle slots, same instance ranges).
//
// A wide node is 128 bytes, 32 words, laid out for one 112-byte fetch of
// seven float4:
//   words  0-3  child min x    words  4-7  child max x
//   words  8-11 child min y    words 12-15 child max y
//   words 16-19 child min z    words 20-23 child max z
//   words 24-27 child entries  words 28-31 unused (zero)
// A child entry is what the traversal pushes on its stack:
//   kWideEmpty                 no child in this position
//   bit 31 clear               interior: index of the child's wide node,
//                              relative to the first node of its tree
//   bit 31 set                 leaf: bits 26-30 hold count - 1 (1..32),
//                              bits 0-25 the first triangle slot (BLAS:
//                              the global slot, index_offset included) or
//                              the first instance_indices entry (TLAS)
// Binary leaves of more than 32 primitives become several leaf entries of at
// most 32 (in order) under extra wide nodes with the leaf's box.
//
// The device reads the quantised form of the same nodes: 64 bytes, 16 words
// (one 64-byte fetch of four uint4):
//   words  0-2  node origin xyz (f32: the minimum corner of the children)
//   word   3    per-axis scale exponents, biased by 127, in bytes 0-2
//               (scale = 2^(byte - 127), a normal float)
//   words  4-9  child planes as bytes (byte k = child k): min x, max x,
//               min y, max y, min z, max z; plane = origin + byte * scale
//   words 10-13 child entries (as above)
//   words 14-15 unused (zero)
// Quantised planes round outwards: every quantised box contains the child's
// box.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pulse_host {

constexpr uint32_t kWideEmpty = 0xffffffffu;
constexpr uint32_t kWideLeafBit = 0x80000000u;
constexpr uint32_t kWideMaxLeaf = 32;
constexpr uint32_t kWideFirstLimit = (1u << 26) - kWideMaxLeaf;  // first + count stays below 2^26
constexpr int kWideNodeWords = 32;

struct WideTree {
    std::vector<uint32_t> words;  // kWideNodeWords per node, root first
    uint32_t stack_need = 0;      // most stack entries a traversal of the tree holds
};

// Collapses the binary tree whose nodes are `nodes[0..n)` (8 words each in
// the reference's node layout: min xyz, first child / first primitive, max
// xyz, primitive count; node 0 is the root, child indices are relative to
// `nodes`) into `out`; leaf entries hold first primitive + `first_offset`.
// False if that is at or beyond kWideFirstLimit (the entry cannot encode it),
// or if an interior entry does not name a later node of `out` (nodes are
// allocated before their children, so this holds for every acyclic input;
// the traversal relies on it instead of counting visits).
bool collapse_wide(const uint32_t *nodes, size_t n, uint32_t first_offset, WideTree &out);

constexpr int kQuantNodeWords = 16;

// Quantised copy of `in` (kQuantNodeWords per node, same node indices and
// entries) appended to `out`.  False if a coordinate is not finite or beyond
// 2^40 in magnitude.
bool quantize_wide(const WideTree &in, std::vector<uint32_t> &out);

// Tighter world boxes for the TLAS leaves of the 4-wide traversal.  The
// reference bounds an instance by the 8 transformed corners of its BLAS
// root box (scene/mod.rs:624-680), which for a rotated mesh is up to 1.7x
// wider per axis than the mesh; bounding the transformed corners of the
// boxes of the BLAS cut at depth kCutDepth (every leaf lies under one of
// them; up to 256 boxes) is far closer.  The binary TLAS keeps the
// reference's boxes.
constexpr int kCutDepth = 8;
// The boxes (min xyz, max xyz) of the BLAS rooted at `nodes` (8 words per
// node, child indices relative to it) cut at depth `depth` (leaves above it
// count as they are), appended to `out`.
void blas_cut_boxes(const uint32_t *nodes, int depth, std::vector<float> &out);
// World box of an instance: the corners of `n` object boxes through the
// column-major object_world `m` (transform_position with the divide by w,
// in f32 as the host builder computes instance boxes), widened outwards by
// 2^-20 of the coordinates' magnitude against that rounding.
void instance_world_box(const float *m, const float *boxes, size_t n, float out[6]);
// instance_world_box for n instances, instance i from matrix m + 16 i (or
// the object_world of pulse_mesh_instance records, stride 40 floats) and its
// cut range, on up to `threads` threads.
void instance_world_boxes(const float *m, size_t m_stride, const float *cut_boxes, const uint32_t *cut_range,
                          size_t n, float *out, int threads);
// Binary tree over n items with boxes boxes[6 j ..] (min xyz, max xyz) in
// the node layout collapse_wide reads (8 words: min xyz, first child or
// first item, max xyz, item count, 0 for interior nodes; children an
// adjacent pair; node 0 the root), one item per leaf.  Splits minimise the
// surface-area cost over the items' box centroids: a full sweep of the
// centroid order along each axis up to 4096 items, 64 centroid bins
// per axis above; coincident centroids split in halves.  The 4-wide
// traversal's TLAS is collapsed from it (the reference's TLAS splits at the
// centroid midpoint, tlas.rs:192).
void sah_tree(const float *boxes, size_t n, std::vector<uint32_t> &nodes);
// The 4-wide traversal's TLAS: sah_tree over the instance_indices entries
// that the leaves of the binary TLAS `tlas` (8 words per node, node 0 the
// root, not the empty-scene sentinel; child and instance ranges checked)
// hold -- the same instances, one leaf each -- bounded by
// inst_box[6 * iidx[entry]], collapsed by collapse_wide.
bool wide_tlas(const uint32_t *tlas, const uint32_t *iidx, const float *inst_box, WideTree &out);
// The nodes of a 4-wide tree grouped by depth, deepest first, for a
// level-by-level refit: out = {n_levels, level_start[0 .. n_levels],
// node indices}; level l holds out[1 + n_levels + 1 + level_start[l] ..
// level_start[l + 1]).  Every child lies in an earlier level than its parent.
void wide_levels(const WideTree &t, std::vector<uint32_t> &out);

// Leaf refinement of the 4-wide BLASes.  The reference's builder stops
// splitting at 8 triangles or fewer (scene/blas.rs:97); the 4-wide
// traversal tests up to two triangles per step, so every binary leaf of
// more than `max_leaf` (and at most kRefineMaxLeaf) triangles becomes a
// binary subtree whose leaves hold at most `max_leaf`, split where the
// surface-area cost (area x triangles, summed over the two sides) over the
// triangles' centroids sorted along one axis is least.  A triangle's box is
// that of p0, p0 + e1, p0 + e2 (the triangle test's own operands) widened
// by 2^-8 of its longest side plus 2^-20 of the coordinates' magnitude,
// so that the points the triangle test accepts just outside the exact
// triangle stay inside.  The triangles of a refined leaf are reordered
// within the leaf's slot range: refinement yields, per global slot g, the
// slot whose triangle sits at g (the identity outside refined leaves).
constexpr uint32_t kRefineMaxLeaf = 64;
constexpr uint32_t kRefineLeaf = 2;  // the upload's max_leaf: a traversal step tests up to two triangles

// Adds, per global slot, the number of leaves of the binary tree `nodes`
// (collapse_wide's form; `n` nodes) that hold it: refs[first_offset +
// first + k] (saturating at 255).  A leaf whose slots other leaves share is
// never refined (its reordering would move their triangles).  With
// `leaf_end`, also stores there the global end slot of the (last) leaf
// holding each slot: of two hits at equal t in one binary leaf the
// reference keeps the lower slot (it tests a leaf in slot order), which
// the 4-wide traversal reproduces from this.
void count_leaf_refs(const uint32_t *nodes, size_t n, uint32_t first_offset, std::vector<uint8_t> &refs,
                     std::vector<uint32_t> *leaf_end = nullptr);

// Copy of the binary tree `nodes` (node 0 the root, `n` nodes) into `out`
// (8 words per node) with each refinable leaf turned into an interior node
// over its appended subtree; `tri` holds 12 floats per global slot (p0, e1,
// e2, 3 unused) and `perm` (one entry per global slot, set to the identity
// beforehand) receives the reordering of the refined leaves' slots.  Leaves
// that any slot of theirs shares with another leaf (refs > 1) stay whole.
void refine_leaves(const uint32_t *nodes, size_t n, uint32_t first_offset, const float *tri, uint32_t max_leaf,
                   const std::vector<uint8_t> &refs, std::vector<uint32_t> &out, uint32_t *perm);

}  // namespace pulse_host
