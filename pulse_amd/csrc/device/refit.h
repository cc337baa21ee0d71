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

// One thread per wide TLAS node: quantise its children's boxes (outwards)
// into words 0-9, as the host's quantize_wide does at upload.
__global__ void __launch_bounds__(64) k_refit_quantize(uint32_t *qnodes, uint32_t n_nodes, const float *child_box,
                                                       uint32_t *err) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n_nodes) return;
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

}  // namespace pt
