// refit.h -- instance motion on the device: new instance transforms, their
// inverses, their world-space boxes and a refit of the quantised 4-wide TLAS
// over them (same tree, new boxes).
//
// The reference rebuilds its TLAS on the host every frame (build_tlas,
// tlas.rs:29-136, from the boxes prepare_mesh_instances computes,
// scene/mod.rs:561-697).  A refit keeps the tree of the last build and only
// recomputes the boxes, so it is exact for the 4-wide traversal (any tree
// over the same instances finds the same closest hits, up to t-ties) but not
// for the reference's binary TLAS, whose node order and shape decide its
// traversal order: that TLAS is left as it was and marked stale.
#pragma once

#include "pt_device.h"

namespace pt {

constexpr uint32_t kErrRefit = 4u;  // a singular transform or a box the quantised nodes cannot hold

// 4x4 inverse by cofactor expansion in double, rounded to f32: the host
// builder's invert4 (scene_builder.cpp), the same operations in the same
// order.  False if singular.
__device__ inline bool invert4_dev(const float *mf, float *outf) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = mf[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) outf[i] = static_cast<float>(inv[i] / det);
    return true;
}

// One wave per instance (block i): lane 0 writes the new object_world and
// its inverse; the lanes share the instance's BLAS cut boxes (cut_range[2 i],
// [2 i + 1] into cut_boxes), transform their corners as the host's
// instance_world_box does and reduce them to the world box, widened by 2^-20
// of its magnitude, in box[6 i .. 6 i + 5].
__global__ void __launch_bounds__(64) k_refit_instances(const float *mats, pulse_mesh_instance *inst,
                                                        const float *cut_boxes, const uint32_t *cut_range,
                                                        float *box, uint32_t n, uint32_t *err) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const uint32_t lane = threadIdx.x;
    float m[16];
    for (int j = 0; j < 16; ++j) m[j] = mats[16u * i + j];
    if (lane == 0u) {
        float wo[16];
        if (invert4_dev(m, wo)) {
            pulse_mesh_instance &mi = inst[i];
            for (int j = 0; j < 16; ++j) {
                mi.object_world[j] = m[j];
                mi.world_object[j] = wo[j];
            }
        } else {
            atomicOr(err, kErrRefit);
        }
    }
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    const uint32_t first = cut_range[2u * i], count = cut_range[2u * i + 1u];
    for (uint32_t j = first + lane; j < first + count; j += 64u) {
        const float *b = cut_boxes + 6u * j;
        for (int c = 0; c < 8; ++c) {
            const float px = (c & 1) ? b[3] : b[0], py = (c & 2) ? b[4] : b[1], pz = (c & 4) ? b[5] : b[2];
            const float hx = m[0] * px + m[4] * py + m[8] * pz + m[12] * 1.0f;
            const float hy = m[1] * px + m[5] * py + m[9] * pz + m[13] * 1.0f;
            const float hz = m[2] * px + m[6] * py + m[10] * pz + m[14] * 1.0f;
            const float hw = m[3] * px + m[7] * py + m[11] * pz + m[15] * 1.0f;
            const float w[3] = {hx / hw, hy / hw, hz / hw};
            for (int a = 0; a < 3; ++a) {
                mn[a] = fminf(mn[a], w[a]);
                mx[a] = fmaxf(mx[a], w[a]);
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
        }
    if (lane == 0u)
        for (int a = 0; a < 3; ++a) {
            box[6u * i + a] = mn[a] - fabsf(mn[a]) * 0x1p-20f;
            box[6u * i + 3u + a] = mx[a] + fabsf(mx[a]) * 0x1p-20f;
        }
}

// The 4-wide traversal's instance records (pt_device.h: kInstRecordWords
// per instance_indices entry) after k_refit_instances: a thread per entry
// rewrites the two matrices of the record and its miss flag from its
// instance (the root, index offset and instance words stay); a bottom row
// other than 0 0 0 1 is an error (the records hold rows 0-2).
__global__ void __launch_bounds__(64) k_refit_records(const pulse_mesh_instance *inst, const uint32_t *iidx,
                                                      uint32_t n_iidx, uint32_t *rec, uint32_t *err) {
    const uint32_t k = blockIdx.x * 64u + threadIdx.x;
    if (k >= n_iidx) return;
    const pulse_mesh_instance &mi = inst[iidx[k]];
    float ow[16], wo[16];
    for (int j = 0; j < 16; ++j) {
        ow[j] = mi.object_world[j];
        wo[j] = mi.world_object[j];
    }
    if (!affine_matrix(ow) || !affine_matrix(wo)) atomicOr(err, kErrRefit);
    uint32_t *r = rec + static_cast<size_t>(kInstRecordWords) * k;
    record_matrix(wo, r + kRecWorldObject);
    r[kRecMissOverflows] = miss_overflows(ow, wo) ? 1u : 0u;
    record_matrix(ow, r + kRecObjectWorld);
}

// One block: the boxes of every child of every wide TLAS node into
// child_box[24 node + 6 k ..] and node_box[6 node ..], one level of the tree
// at a time, deepest first (`levels`: {n_levels, level starts, node
// indices}, host/wide_bvh.h: wide_levels), a thread per node of a level; a
// node's children lie in earlier levels, done before the barrier.
__global__ void __launch_bounds__(256) k_refit_boxes(const uint32_t *qnodes, const uint32_t *levels,
                                                     const uint32_t *iidx, const float *inst_box, float *child_box,
                                                     float *node_box) {
    const uint32_t n_levels = levels[0];
    const uint32_t *order = levels + 2u + n_levels;
    for (uint32_t l = 0; l < n_levels; ++l) {
        for (uint32_t j = levels[1u + l] + threadIdx.x; j < levels[2u + l]; j += blockDim.x) {
            const uint32_t i = order[j];
            float nb[6] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f,
                           -3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
            for (int k = 0; k < 4; ++k) {
                const uint32_t e = qnodes[16u * i + 10u + k];
                float cb[6] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f,
                               -3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
                if (e == 0xffffffffu) {
                    // empty position
                } else if (e >> 31) {
                    const uint32_t first = e & ((1u << 26) - 1u), count = ((e >> 26) & 31u) + 1u;
                    for (uint32_t m = first; m < first + count; ++m) {
                        const float *b = inst_box + 6u * iidx[m];
                        for (int a = 0; a < 3; ++a) {
                            cb[a] = fminf(cb[a], b[a]);
                            cb[3 + a] = fmaxf(cb[3 + a], b[3 + a]);
                        }
                    }
                } else {
                    for (int a = 0; a < 6; ++a) cb[a] = node_box[6u * e + a];
                }
                for (int a = 0; a < 6; ++a) child_box[24u * i + 6u * k + a] = cb[a];
                if (e != 0xffffffffu)
                    for (int a = 0; a < 3; ++a) {
                        nb[a] = fminf(nb[a], cb[a]);
                        nb[3 + a] = fmaxf(nb[3 + a], cb[3 + a]);
                    }
            }
            for (int a = 0; a < 6; ++a) node_box[6u * i + a] = nb[a];
        }
        __syncthreads();
    }
}

// Quantises wide TLAS node i's children's boxes (child_box[24 i + 6 k ..])
// outwards into words 0-9 of its 16, as the host's quantize_wide does at
// upload; a box the node cannot hold raises kErrRefit.
__device__ inline void quantize_tlas_node(uint32_t *qnodes, uint32_t i, const float *child_box, uint32_t *err) {
    uint32_t *q = qnodes + 16u * i;
    bool valid[4];
    for (int k = 0; k < 4; ++k) valid[k] = q[10 + k] != 0xffffffffu;
    uint32_t exps = 0, planes[6] = {0, 0, 0, 0, 0, 0};
    float origin[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; ++a) {
        float lo = 0.0f, hi = 0.0f;
        bool any = false;
        for (int k = 0; k < 4; ++k) {
            if (!valid[k]) continue;
            const float cl = child_box[24u * i + 6u * k + a], ch = child_box[24u * i + 6u * k + 3u + a];
            if (!isfinite(cl) || !isfinite(ch) || fabsf(cl) > 0x1p40f || fabsf(ch) > 0x1p40f) {
                atomicOr(err, kErrRefit);
                return;
            }
            lo = any ? fminf(lo, cl) : cl;
            hi = any ? fmaxf(hi, ch) : ch;
            any = true;
        }
        const double extent = static_cast<double>(hi) - lo;
        int e = -126;
        if (extent > 0.0) {
            e = max(-126, static_cast<int>(ceil(log2(extent / 255.0))));
            while (extent > 255.0 * ldexp(1.0, e)) ++e;
        }
        if (e > 127) {
            atomicOr(err, kErrRefit);
            return;
        }
        const double sc = ldexp(1.0, e);
        exps |= static_cast<uint32_t>(e + 127) << (8 * a);
        origin[a] = lo;
        for (int k = 0; k < 4; ++k) {
            if (!valid[k]) continue;
            const float cl = child_box[24u * i + 6u * k + a], ch = child_box[24u * i + 6u * k + 3u + a];
            double l = floor((static_cast<double>(cl) - lo) / sc);
            double h = ceil((static_cast<double>(ch) - lo) / sc);
            while (l > 0.0 && lo + l * sc > cl) l -= 1.0;
            while (h < 255.0 && lo + h * sc < ch) h += 1.0;
            if (l < 0.0 || h > 255.0 || lo + l * sc > cl || lo + h * sc < ch) {
                atomicOr(err, kErrRefit);
                return;
            }
            planes[2 * a] |= static_cast<uint32_t>(l) << (8 * k);
            planes[2 * a + 1] |= static_cast<uint32_t>(h) << (8 * k);
        }
    }
    for (int a = 0; a < 3; ++a) q[a] = __float_as_uint(origin[a]);
    q[3] = exps;
    for (int j = 0; j < 6; ++j) q[4 + j] = planes[j];
}

// One thread per wide TLAS node: quantize_tlas_node.
__global__ void __launch_bounds__(64) k_refit_quantize(uint32_t *qnodes, uint32_t n_nodes, const float *child_box,
                                                       uint32_t *err) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i < n_nodes) quantize_tlas_node(qnodes, i, child_box, err);
}

// ---------------------------------------------------------------------------
// Rebuild of the 4-wide TLAS on the device (pulse_scene_rebuild_instances).
// A refit keeps the tree of the last build: after large motion its boxes
// overlap and traversal slows (all 800 instances of the bench scene moved to
// new places: 12.9 ms per frame refit against 5.3 ms rebuilt).  One block
// builds a new tree over the instance_indices entries from the instances'
// world boxes, in LDS and a scratch array:
//   1. 64-bit keys: a 30-bit Morton code of the box centre within the
//      centres' bounds, then the entry index (keys are distinct); bitonic
//      sort in LDS;
//   2. the binary radix tree over the sorted keys (Karras 2012, "Maximizing
//      parallelism in the construction of BVHs, octrees, and k-d trees"):
//      internal node i covers a key range and splits at the highest
//      differing bit; node 0 is the root; every leaf is one entry;
//   3. boxes bottom up (a counter per internal node: the second child to
//      arrive computes the node), depths by walking to the root;
//   4. wide nodes = the internal nodes at even depth, each over its
//      grandchildren (a leaf child stays a leaf), numbered level by level
//      (a second sort by (depth / 2, node)), so every interior entry names a
//      later node and the traversal, which counts no visits, ends;
//   5. quantised as at upload; the level table for later refits.
// Every loop that follows parent links stops after n steps (a malformed
// tree raises kErrRefit instead of looping).
constexpr uint32_t kRebuildMax = 4096;   // entries one block sorts in LDS
constexpr int kRebuildThreads = 1024;
constexpr uint32_t kRbNone = 0xffffffffu;
constexpr uint32_t kRbLeaf = 0x80000000u;
constexpr uint32_t kRebuildScratchWords = 16;  // per entry, see k_rebuild_tlas

__device__ __forceinline__ uint32_t morton_expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ int rb_delta(const uint64_t *keys, int n, int a, int b) {
    if (b < 0 || b > n - 1) return -1;
    return __clzll(static_cast<long long>(keys[a] ^ keys[b]));
}
// Ascending bitonic sort of keys[0 .. p), p a power of two, by the block.
__device__ inline void rb_bitonic(uint64_t *keys, uint32_t p) {
    for (uint32_t k = 2; k <= p; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = keys[i], b = keys[ixj];
                    const bool up = (i & k) == 0u;
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// n entries (1 .. kRebuildMax), entry j's box inst_box[6 iidx[j] ..];
// scratch holds kRebuildScratchWords n words; qnodes room for max(n - 1, 1)
// nodes, child_box 24 floats per node, levels 2 + 2 max(n - 1, 1) words.
// info[0] = wide nodes, info[1] = wide levels.
__global__ void __launch_bounds__(kRebuildThreads) k_rebuild_tlas(const uint32_t *iidx, uint32_t n,
                                                                   const float *inst_box, uint32_t *scratch,
                                                                   float *child_box, uint32_t *qnodes,
                                                                   uint32_t *levels, uint32_t *info, uint32_t *err) {
    __shared__ uint64_t keys[kRebuildMax];
    __shared__ float red[6][kRebuildThreads / 64];
    __shared__ float bounds[6];
    __shared__ uint32_t offs[72];
    __shared__ uint32_t counts[2];
    const uint32_t tid = threadIdx.x, T = blockDim.x;
    const uint32_t N = n, M = n - 1u;  // leaves, internal nodes
    uint32_t *pos = scratch;           // N: sorted leaf -> entry
    uint32_t *left = pos + N;          // M each below
    uint32_t *right = left + M;
    uint32_t *par_i = right + M;
    uint32_t *par_l = par_i + M;       // N
    uint32_t *flag = par_l + N;
    uint32_t *depth = flag + M;
    uint32_t *widx = depth + M;
    float *ibox = reinterpret_cast<float *>(widx + M);  // 6 M
    auto box_of = [&](uint32_t c, float *b) {
        const float *src = (c & kRbLeaf) ? inst_box + 6u * iidx[pos[c & ~kRbLeaf]] : ibox + 6u * c;
        for (int a = 0; a < 6; ++a) b[a] = src[a];
    };
    if (N == 1u) {
        if (tid == 0) {
            const float *b = inst_box + 6u * iidx[0];
            uint32_t *q = qnodes;
            for (int k = 0; k < 16; ++k) q[k] = 0u;
            q[10] = kRbLeaf | 0u;
            q[11] = q[12] = q[13] = kRbNone;
            for (int a = 0; a < 24; ++a) child_box[a] = 0.0f;
            for (int a = 0; a < 6; ++a) child_box[a] = b[a];
            quantize_tlas_node(qnodes, 0u, child_box, err);
            levels[0] = 1u;
            levels[1] = 0u;
            levels[2] = 1u;
            levels[3] = 0u;
            info[0] = 1u;
            info[1] = 1u;
        }
        return;
    }
    // 1. centre bounds, keys, sort
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (uint32_t j = tid; j < N; j += T) {
        const float *b = inst_box + 6u * iidx[j];
        for (int a = 0; a < 3; ++a) {
            const float c = 0.5f * b[a] + 0.5f * b[3 + a];
            mn[a] = fminf(mn[a], c);
            mx[a] = fmaxf(mx[a], c);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
        }
    if ((tid & 63u) == 0u)
        for (int a = 0; a < 3; ++a) {
            red[a][tid >> 6] = mn[a];
            red[3 + a][tid >> 6] = mx[a];
        }
    __syncthreads();
    if (tid < 6u) {
        float v = red[tid][0];
        for (uint32_t w = 1; w < T / 64u; ++w) v = tid < 3u ? fminf(v, red[tid][w]) : fmaxf(v, red[tid][w]);
        bounds[tid] = v;
    }
    __syncthreads();
    uint32_t P = 1;
    while (P < N) P <<= 1;
    for (uint32_t j = tid; j < P; j += T) {
        uint64_t key = ~0ull;
        if (j < N) {
            const float *b = inst_box + 6u * iidx[j];
            uint32_t q[3];
            for (int a = 0; a < 3; ++a) {
                const float ext = bounds[3 + a] - bounds[a];
                const float c = 0.5f * b[a] + 0.5f * b[3 + a];
                const float u = ext > 0.0f ? (c - bounds[a]) / ext * 1024.0f : 0.0f;
                q[a] = static_cast<uint32_t>(fminf(fmaxf(u, 0.0f), 1023.0f));
            }
            const uint32_t m = (morton_expand10(q[0]) << 2) | (morton_expand10(q[1]) << 1) | morton_expand10(q[2]);
            key = (static_cast<uint64_t>(m) << 32) | j;
        }
        keys[j] = key;
    }
    __syncthreads();
    rb_bitonic(keys, P);
    // 2. radix tree
    for (uint32_t i = tid; i < N; i += T) pos[i] = static_cast<uint32_t>(keys[i]);
    if (tid == 0) par_i[0] = kRbNone;
    for (uint32_t i = tid; i < M; i += T) {
        const int ii = static_cast<int>(i), nn = static_cast<int>(N);
        const int d = rb_delta(keys, nn, ii, ii + 1) - rb_delta(keys, nn, ii, ii - 1) > 0 ? 1 : -1;
        const int dmin = rb_delta(keys, nn, ii, ii - d);
        int lmax = 2;
        while (rb_delta(keys, nn, ii, ii + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (rb_delta(keys, nn, ii, ii + (l + t) * d) > dmin) l += t;
        const int j = ii + l * d;
        const int dnode = rb_delta(keys, nn, ii, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (rb_delta(keys, nn, ii, ii + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = ii + s * d + min(d, 0);
        const uint32_t L = min(ii, j) == gamma ? (kRbLeaf | static_cast<uint32_t>(gamma)) : static_cast<uint32_t>(gamma);
        const uint32_t R = max(ii, j) == gamma + 1 ? (kRbLeaf | static_cast<uint32_t>(gamma + 1))
                                                   : static_cast<uint32_t>(gamma + 1);
        left[i] = L;
        right[i] = R;
        if (L & kRbLeaf) par_l[L & ~kRbLeaf] = i;
        else par_i[L] = i;
        if (R & kRbLeaf) par_l[R & ~kRbLeaf] = i;
        else par_i[R] = i;
        flag[i] = 0u;
    }
    __syncthreads();
    // 3. boxes bottom up, depths
    for (uint32_t i = tid; i < N; i += T) {
        uint32_t p = par_l[i];
        for (uint32_t steps = 0; p != kRbNone; ++steps) {
            if (steps > N || p >= M) {
                atomicOr(err, kErrRefit);
                break;
            }
            __threadfence();
            if (atomicAdd(&flag[p], 1u) == 0u) break;  // the other child completes p
            __threadfence();
            float a[6], b[6];
            box_of(left[p], a);
            box_of(right[p], b);
            for (int k = 0; k < 3; ++k) {
                ibox[6u * p + k] = fminf(a[k], b[k]);
                ibox[6u * p + 3 + k] = fmaxf(a[3 + k], b[3 + k]);
            }
            p = par_i[p];
        }
    }
    for (uint32_t i = tid; i < M; i += T) {
        uint32_t dd = 0, p = i;
        while (p != 0u && dd <= M) {
            p = par_i[p];
            if (p >= M) break;
            ++dd;
        }
        if (p != 0u || dd > M) atomicOr(err, kErrRefit);
        depth[i] = dd;
    }
    __syncthreads();
    // 4. wide nodes, numbered level by level
    uint32_t P2 = 1;
    while (P2 < M) P2 <<= 1;
    for (uint32_t i = tid; i < P2; i += T)
        keys[i] = (i < M && (depth[i] & 1u) == 0u) ? ((static_cast<uint64_t>(depth[i] >> 1) << 32) | i) : ~0ull;
    __syncthreads();
    rb_bitonic(keys, P2);
    for (uint32_t k = tid; k < P2; k += T)
        if (keys[k] != ~0ull && (k + 1u == P2 || keys[k + 1u] == ~0ull)) {
            counts[0] = k + 1u;                                          // wide nodes
            counts[1] = static_cast<uint32_t>(keys[k] >> 32) + 1u;       // wide levels
        }
    __syncthreads();
    const uint32_t n_wide = counts[0], n_levels = counts[1];
    if (n_levels > 70u) {  // deeper than 64-bit keys allow: malformed
        if (tid == 0) atomicOr(err, kErrRefit);
        return;
    }
    for (uint32_t k = tid; k < n_wide; k += T) {
        widx[static_cast<uint32_t>(keys[k])] = k;
        const uint32_t lv = static_cast<uint32_t>(keys[k] >> 32);
        if (k == 0u || static_cast<uint32_t>(keys[k - 1u] >> 32) != lv) offs[lv] = k;
    }
    if (tid == 0) offs[n_levels] = n_wide;
    __syncthreads();
    for (uint32_t k = tid; k < n_wide; k += T) {
        const uint32_t b = static_cast<uint32_t>(keys[k]);
        uint32_t *q = qnodes + 16u * k;
        float *cb = child_box + 24u * k;
        int slot = 0;
        auto put = [&](uint32_t c) {
            float bx[6];
            box_of(c, bx);
            uint32_t e;
            if (c & kRbLeaf) {
                e = kRbLeaf | pos[c & ~kRbLeaf];  // one entry, count 1
            } else {
                e = widx[c];
                if (e <= k || e >= n_wide) atomicOr(err, kErrRefit);
            }
            q[10 + slot] = e;
            for (int a = 0; a < 6; ++a) cb[6 * slot + a] = bx[a];
            ++slot;
        };
        for (int h = 0; h < 2; ++h) {
            const uint32_t c = h == 0 ? left[b] : right[b];
            if (c & kRbLeaf) {
                put(c);
            } else {
                put(left[c]);
                put(right[c]);
            }
        }
        for (; slot < 4; ++slot) {
            q[10 + slot] = kRbNone;
            for (int a = 0; a < 6; ++a) cb[6 * slot + a] = 0.0f;
        }
        q[14] = 0u;
        q[15] = 0u;
        quantize_tlas_node(qnodes, k, child_box, err);
        // level table, deepest level first
        const uint32_t lv = static_cast<uint32_t>(keys[k] >> 32);
        levels[2u + n_levels + (n_wide - offs[lv + 1u]) + (k - offs[lv])] = k;
    }
    for (uint32_t l = tid; l <= n_levels; l += T) levels[1u + l] = n_wide - offs[n_levels - l];
    if (tid == 0) {
        levels[0] = n_levels;
        info[0] = n_wide;
        info[1] = n_levels;
    }
}

}  // namespace pt
